# oracle/_ref: the reference's own leaf modules, compiled in place from the
# reference tree (no copies, no stand-ins): constants, the Sedov setup's
# setup_config_mod (kappa = 7/5), Euler equations, Riemann solvers,
# two-point fluxes, RK constants and interpolation utilities.
REF_BASE := $(REF)/utils/constants_mod.f90 \
            $(REF)/setups/euler/3d/sedov/setup_config_mod.f90 \
            $(REF)/equations/euler/equations_const_mod.f90 \
            $(REF)/equations/euler/equations_mod.f90
REF_A := $(REF_BASE) \
         $(REF)/equations/euler/riemann/hllc/riemann_mod.f90 \
         $(REF)/equations/euler/two_point/chandrashekar/two_point_flux_mod.f90 \
         $(REF)/timedisc/rungekutta/low-storage/4-5/timedisc_const_mod.f90 \
         $(REF)/utils/linalg_mod.f90 \
         $(REF)/utils/interpol_mod.f90 \
         $(REF)/utils/math_mod.f90 \
         ref_shim_a.f90
REF_B := $(REF_BASE) \
         $(REF)/equations/euler/riemann/rusanov/riemann_mod.f90 \
         $(REF)/equations/euler/two_point/standard/two_point_flux_mod.f90 \
         $(REF)/timedisc/rungekutta/ssp/4-5/timedisc_const_mod.f90 \
         ref_shim_b.f90

_ref/libnemo_ref_a.so: ref_shim_a.f90 ref.mk
	rm -rf _ref/a && mkdir -p _ref/a
	cd _ref/a && for f in $(REF_A); do \
	  case $$f in /*) src=$$f;; *) src=../../$$f;; esac; \
	  $(FC) $(FFLAGS) -c $$src -o $$(basename $$f .f90).o || exit 1; done
	$(FC) -shared -Wl,-Bsymbolic -o $@ _ref/a/*.o

_ref/libnemo_ref_b.so: ref_shim_b.f90 ref.mk
	rm -rf _ref/b && mkdir -p _ref/b
	cd _ref/b && for f in $(REF_B); do \
	  case $$f in /*) src=$$f;; *) src=../../$$f;; esac; \
	  $(FC) $(FFLAGS) -c $$src -o $$(basename $$f .f90).o || exit 1; done
	$(FC) -shared -Wl,-Bsymbolic -o $@ _ref/b/*.o

# Library C: GLM-MHD (PP_N_VARS=9) with the Orszag-Tang setup's constants.
FFLAGS_MHD := -cpp -O2 -fPIC -ffp-contract=off -DPP_N_NODES=4 -DPP_N_DIMS=3 -DPP_N_VARS=9 -DPP_EQN_MHD=1
REF_C := $(REF)/utils/constants_mod.f90 \
         $(REF)/setups/mhd/3d/orszag-tang/setup_config_mod.f90 \
         $(REF)/equations/mhd/equations_const_mod.f90 \
         $(REF)/equations/mhd/equations_share_mod.f90 \
         $(REF)/equations/mhd/equations_mod.f90 \
         $(REF)/equations/mhd/riemann/rusanov/riemann_mod.f90 \
         $(REF)/equations/mhd/non_conservative_flux_mod.f90 \
         ref_shim_c.f90

_ref/libnemo_ref_c.so: ref_shim_c.f90 ref.mk
	rm -rf _ref/c && mkdir -p _ref/c
	cd _ref/c && for f in $(REF_C); do \
	  case $$f in /*) src=$$f;; *) src=../../$$f;; esac; \
	  $(FC) $(FFLAGS_MHD) -c $$src -o $$(basename $$f .f90).o 2>/dev/null || exit 1; done
	$(FC) -shared -Wl,-Bsymbolic -o $@ _ref/c/*.o
