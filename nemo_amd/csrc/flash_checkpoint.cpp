// flash_checkpoint.cpp -- FLASH-format HDF5 checkpoint writer (host side).
//
// Writes the file layout of the reference's checkpoint module
// (checkpoint/flash/checkpoint_mod.f90:11-364, checkpoint_utils_mod.f90,
// datasets of setups/euler/3d/_common_/setup_checkpoint_mod.f90:9-60) with
// the HDF5 C library (1.10, the version the reference builds against):
// "sim info", the name/value tables of runtime parameters and scalars,
// "unknown names", the block tree (coordinates, block size, bounding box,
// gid, node type, refine level, which child, bflags) and one dataset per
// variable of shape [numblocks][nzb][nyb][nxb], leaves last (parents stay
// zero, as the reference leaves them).  The file is written under
// <path>.partial and renamed when complete (checkpoint_mod.f90:228,359).
// The Fortran arrays (3, numblocks) etc. appear with their dimensions
// reversed, as HDF5 stores Fortran arrays.
#include <hdf5.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nemo_flash.h"

namespace {

std::string& last_error() {
    static thread_local std::string msg;
    return msg;
}

int fail(const std::string& msg) {
    last_error() = msg;
    return NEMO_FLASH_ERR;
}

// Owns an HDF5 identifier; closes it with the matching routine.
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Hid(const Hid&) = delete;
    Hid& operator=(const Hid&) = delete;
    ~Hid() {
        if (id >= 0 && close) close(id);
    }
    bool ok() const { return id >= 0; }
};

hid_t str_type(size_t n) {
    hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, n);
    return t;
}

void put_str(char* dst, size_t n, const std::string& s) {
    std::memset(dst, 0, n);
    std::memcpy(dst, s.data(), s.size() < n ? s.size() : n);
}

// One-dimensional dataset of {name: string(80), value: T}
// (checkpoint_write_*_runtime_parameters / _scalars).
template <typename T>
int write_table(hid_t file, const char* dname, hid_t value_type, const std::vector<std::pair<std::string, T>>& rows) {
    struct Row {
        char name[80];
        T value;
    };
    Hid s80(str_type(80), H5Tclose);
    Hid ct(H5Tcreate(H5T_COMPOUND, sizeof(Row)), H5Tclose);
    H5Tinsert(ct.id, "name", HOFFSET(Row, name), s80.id);
    H5Tinsert(ct.id, "value", HOFFSET(Row, value), value_type);
    std::vector<Row> buf(rows.size());
    for (size_t i = 0; i < rows.size(); ++i) {
        put_str(buf[i].name, 80, rows[i].first);
        buf[i].value = rows[i].second;
    }
    hsize_t dims[1] = {hsize_t(rows.size())};
    Hid sp(H5Screate_simple(1, dims, nullptr), H5Sclose);
    Hid ds(H5Dcreate2(file, dname, ct.id, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!ds.ok() || H5Dwrite(ds.id, ct.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
        return fail(std::string("cannot write ") + dname);
    return NEMO_FLASH_OK;
}

int write_strtable(hid_t file, const char* dname, const std::vector<std::pair<std::string, std::string>>& rows) {
    struct Row {
        char name[80];
        char value[80];
    };
    Hid s80(str_type(80), H5Tclose);
    Hid ct(H5Tcreate(H5T_COMPOUND, sizeof(Row)), H5Tclose);
    H5Tinsert(ct.id, "name", HOFFSET(Row, name), s80.id);
    H5Tinsert(ct.id, "value", HOFFSET(Row, value), s80.id);
    std::vector<Row> buf(rows.size());
    for (size_t i = 0; i < rows.size(); ++i) {
        put_str(buf[i].name, 80, rows[i].first);
        put_str(buf[i].value, 80, rows[i].second);
    }
    hsize_t dims[1] = {hsize_t(rows.size())};
    Hid sp(H5Screate_simple(1, dims, nullptr), H5Sclose);
    Hid ds(H5Dcreate2(file, dname, ct.id, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!ds.ok() || H5Dwrite(ds.id, ct.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
        return fail(std::string("cannot write ") + dname);
    return NEMO_FLASH_OK;
}

// Plain dataset of a native type; rank and C-order dims as given.
int write_array(hid_t file, const char* dname, hid_t type, int rank, const hsize_t* dims, const void* data) {
    Hid lcpl(H5Pcreate(H5P_LINK_CREATE), H5Pclose);
    H5Pset_create_intermediate_group(lcpl.id, 1);
    Hid dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    H5Pset_alloc_time(dcpl.id, H5D_ALLOC_TIME_EARLY);  // unwritten parts read as zero (parent blocks)
    Hid sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
    Hid ds(H5Dcreate2(file, dname, type, sp.id, lcpl.id, dcpl.id, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(std::string("cannot create ") + dname);
    if (data && H5Dwrite(ds.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
        return fail(std::string("cannot write ") + dname);
    return NEMO_FLASH_OK;
}

// Leaf rows [first, first + count) of a [numblocks][...] dataset.
int write_rows(hid_t file, const char* dname, hid_t type, int rank, const hsize_t* dims, hsize_t first,
               hsize_t count, const void* data) {
    Hid lcpl(H5Pcreate(H5P_LINK_CREATE), H5Pclose);
    H5Pset_create_intermediate_group(lcpl.id, 1);
    Hid dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    H5Pset_alloc_time(dcpl.id, H5D_ALLOC_TIME_EARLY);
    Hid fsp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
    Hid ds(H5Dcreate2(file, dname, type, fsp.id, lcpl.id, dcpl.id, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(std::string("cannot create ") + dname);
    if (count == 0) return NEMO_FLASH_OK;
    std::vector<hsize_t> start(rank, 0), cnt(dims, dims + rank);
    start[0] = first;
    cnt[0] = count;
    H5Sselect_hyperslab(fsp.id, H5S_SELECT_SET, start.data(), nullptr, cnt.data(), nullptr);
    Hid msp(H5Screate_simple(rank, cnt.data(), nullptr), H5Sclose);
    if (H5Dwrite(ds.id, type, msp.id, fsp.id, H5P_DEFAULT, data) < 0)
        return fail(std::string("cannot write ") + dname);
    return NEMO_FLASH_OK;
}

int write_file(const char* path, const nemo_flash_meta* m, const nemo_flash_tree* t, int32_t nunk,
               const char* names, const double* leaf_data, const double* leaf_blend) {
    Hid file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) return fail(std::string("cannot create ") + path);
    const hid_t f = file.id;

    {  // sim info (checkpoint_write_sim_info, checkpoint_utils_mod.f90:399-585)
        struct SimInfo {
            int file_format_version;
            char setup_call[400], file_creation_time[80], flash_version[80], build_date[80], build_dir[80],
                build_machine[80], cflags[400], fflags[400], setup_time_stamp[80], build_time_stamp[80];
        } si;
        std::memset(&si, 0, sizeof(si));
        si.file_format_version = m->file_format_version;
        put_str(si.setup_call, 400, "none");
        put_str(si.file_creation_time, 80, "none");
        put_str(si.flash_version, 80, "FLASH 4.3_release");
        put_str(si.build_date, 80, "none");
        put_str(si.build_dir, 80, "none");
        put_str(si.build_machine, 80, "none");
        put_str(si.cflags, 400, "none");
        put_str(si.fflags, 400, "none");
        put_str(si.setup_time_stamp, 80, "none");
        put_str(si.build_time_stamp, 80, "none");
        Hid s80(str_type(80), H5Tclose), s400(str_type(400), H5Tclose);
        Hid ct(H5Tcreate(H5T_COMPOUND, sizeof(SimInfo)), H5Tclose);
        H5Tinsert(ct.id, "file format version", HOFFSET(SimInfo, file_format_version), H5T_NATIVE_INT);
        H5Tinsert(ct.id, "setup call", HOFFSET(SimInfo, setup_call), s400.id);
        H5Tinsert(ct.id, "file creation time", HOFFSET(SimInfo, file_creation_time), s80.id);
        H5Tinsert(ct.id, "flash version", HOFFSET(SimInfo, flash_version), s80.id);
        H5Tinsert(ct.id, "build date", HOFFSET(SimInfo, build_date), s80.id);
        H5Tinsert(ct.id, "build dir", HOFFSET(SimInfo, build_dir), s80.id);
        H5Tinsert(ct.id, "build machine", HOFFSET(SimInfo, build_machine), s80.id);
        H5Tinsert(ct.id, "cflags", HOFFSET(SimInfo, cflags), s400.id);
        H5Tinsert(ct.id, "fflags", HOFFSET(SimInfo, fflags), s400.id);
        H5Tinsert(ct.id, "setup time stamp", HOFFSET(SimInfo, setup_time_stamp), s80.id);
        H5Tinsert(ct.id, "build time stamp", HOFFSET(SimInfo, build_time_stamp), s80.id);
        hsize_t one[1] = {1};
        Hid sp(H5Screate_simple(1, one, nullptr), H5Sclose);
        Hid ds(H5Dcreate2(f, "sim info", ct.id, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
        if (!ds.ok() || H5Dwrite(ds.id, ct.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, &si) < 0)
            return fail("cannot write sim info");
    }

    // runtime parameters and scalars (hook_checkpoint_200, checkpoint_mod.f90:43-156)
    std::vector<std::pair<std::string, int>> ipar = {
        {"lrefine_min", m->lrefine_min}, {"lrefine_max", m->lrefine_max},
        {"checkpointfilenumber", m->file_number}, {"checkpointfileintervalstep", 0},
        {"fileformatversion", m->file_format_version}, {"forcedplotfilenumber", 0}};
    std::vector<std::pair<std::string, double>> rpar = {
        {"xmin", m->xmin}, {"xmax", m->xmax}, {"ymin", m->ymin}, {"ymax", m->ymax}, {"zmin", m->zmin},
        {"zmax", m->zmax}, {"cfl", m->cfl}, {"checkpointfileintervaltime", m->checkpoint_interval_time},
        {"tinitial", m->tinitial}, {"tmax", m->tmax}, {"gamma", m->gamma}};
    std::vector<std::pair<std::string, int>> lpar = {{"restart", 0}};
    std::vector<std::pair<std::string, std::string>> spar = {
        {"unitsystem", "none"}, {"geometry", "cartesian"}, {"basen", "chkpt_"}};
    const char* const bkey[6] = {"xl_boundary_type", "xr_boundary_type", "yl_boundary_type",
                                 "yr_boundary_type", "zl_boundary_type", "zr_boundary_type"};
    const char* const bname[3] = {"periodic", "outflow", "reflecting"};
    for (int side = 0; side < 6; ++side) spar.emplace_back(bkey[side], bname[int(m->boundary[side])]);
    std::vector<std::pair<std::string, int>> iscal = {
        {"nstep", m->nstep}, {"nxb", m->nxb}, {"nyb", m->nxb}, {"nzb", m->nxb}, {"dimensionality", 3},
        {"globalnumblocks", int(t->numblocks)}, {"checkpointfilenumber", m->file_number}};
    std::vector<std::pair<std::string, double>> rscal = {
        {"time", m->time}, {"dt", m->dt}, {"dtold", m->dt}, {"dtnew", m->dt},
        {"nextcheckpointtime", m->next_checkpoint_time}};
    std::vector<std::pair<std::string, int>> lscal = {{"doubleprecision", 1}};
    std::vector<std::pair<std::string, std::string>> sscal = {{"geometry", "cartesian"}, {"nemo", "test"}};
    int rc;
    if ((rc = write_table(f, "real runtime parameters", H5T_NATIVE_DOUBLE, rpar))) return rc;
    if ((rc = write_strtable(f, "string runtime parameters", spar))) return rc;
    if ((rc = write_table(f, "logical runtime parameters", H5T_NATIVE_INT, lpar))) return rc;
    if ((rc = write_table(f, "integer runtime parameters", H5T_NATIVE_INT, ipar))) return rc;
    if ((rc = write_table(f, "real scalars", H5T_NATIVE_DOUBLE, rscal))) return rc;
    if ((rc = write_strtable(f, "string scalars", sscal))) return rc;
    if ((rc = write_table(f, "logical scalars", H5T_NATIVE_INT, lscal))) return rc;
    if ((rc = write_table(f, "integer scalars", H5T_NATIVE_INT, iscal))) return rc;

    {  // unknown names: string(4), Fortran shape (1, nunk)
        std::vector<char> buf(size_t(nunk) * 4, 0);
        std::memcpy(buf.data(), names, buf.size());
        Hid s4(str_type(4), H5Tclose);
        hsize_t dims[2] = {hsize_t(nunk), 1};
        if ((rc = write_array(f, "unknown names", s4.id, 2, dims, buf.data()))) return rc;
    }

    // block tree (hook_checkpoint_300, checkpoint_mod.f90:245-300)
    const hsize_t nb = hsize_t(t->numblocks);
    hsize_t d3[2] = {nb, 3}, dbb[3] = {nb, 3, 2}, dgid[2] = {nb, 15}, d1[1] = {nb};
    if ((rc = write_array(f, "coordinates", H5T_NATIVE_DOUBLE, 2, d3, t->coordinates))) return rc;
    if ((rc = write_array(f, "block size", H5T_NATIVE_DOUBLE, 2, d3, t->block_size))) return rc;
    if ((rc = write_array(f, "bounding box", H5T_NATIVE_DOUBLE, 3, dbb, t->bounding_box))) return rc;
    if ((rc = write_array(f, "gid", H5T_NATIVE_INT, 2, dgid, t->gid))) return rc;
    if ((rc = write_array(f, "node type", H5T_NATIVE_INT, 1, d1, t->node_type))) return rc;
    if ((rc = write_array(f, "refine level", H5T_NATIVE_INT, 1, d1, t->refine_level))) return rc;
    if ((rc = write_array(f, "which child", H5T_NATIVE_INT, 1, d1, t->which_child))) return rc;
    std::vector<int> bflags(nb, -1);
    if ((rc = write_array(f, "bflags", H5T_NATIVE_INT, 1, d1, bflags.data()))) return rc;

    // variables (setup_checkpoint_mod.f90:31-60): leaves occupy the last rows
    const hsize_t nl = hsize_t(t->nleaves), n3 = hsize_t(m->nxb) * m->nxb * m->nxb;
    const hsize_t first = nb - nl;
    hsize_t dv[4] = {nb, hsize_t(m->nxb), hsize_t(m->nxb), hsize_t(m->nxb)};
    for (int v = 0; v < nunk; ++v) {
        std::string dn = "/" + std::string(names + 4 * v, strnlen(names + 4 * v, 4));
        if ((rc = write_rows(f, dn.c_str(), H5T_NATIVE_DOUBLE, 4, dv, first, nl, leaf_data + size_t(v) * nl * n3)))
            return rc;
    }
    if (leaf_blend) {
        if ((rc = write_rows(f, "/blnd", H5T_NATIVE_DOUBLE, 4, dv, first, nl, leaf_blend))) return rc;
    }
    if (t->dtmin) {
        std::vector<int> rank(nb, 0), minloc(nb, 0);
        if ((rc = write_array(f, "/mesh/mpirank", H5T_NATIVE_INT, 1, d1, rank.data()))) return rc;
        if ((rc = write_rows(f, "/minloc", H5T_NATIVE_INT, 1, d1, first, nl, minloc.data()))) return rc;
        if ((rc = write_rows(f, "/dtmin", H5T_NATIVE_DOUBLE, 1, d1, first, nl, t->dtmin))) return rc;
    }
    return NEMO_FLASH_OK;
}

}  // namespace

extern "C" {

int nemo_flash_write(const char* path, const nemo_flash_meta* meta, const nemo_flash_tree* tree, int32_t nunk,
                     const char* unknown_names, const double* leaf_data, const double* leaf_blend) {
    if (!path || !meta || !tree || nunk <= 0 || !unknown_names || !leaf_data || !tree->coordinates ||
        !tree->block_size || !tree->bounding_box || !tree->gid || !tree->node_type || !tree->refine_level ||
        !tree->which_child || tree->numblocks <= 0 || tree->nleaves <= 0 || tree->nleaves > tree->numblocks ||
        meta->nxb <= 0)
        return fail("invalid checkpoint arguments");
    for (int side = 0; side < 6; ++side)
        if (meta->boundary[side] < 0 || meta->boundary[side] > 2)
            return fail("boundary type of a box side out of range");
    const std::string tmp = std::string(path) + ".partial";
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    int rc = write_file(tmp.c_str(), meta, tree, nunk, unknown_names, leaf_data, leaf_blend);
    if (rc != NEMO_FLASH_OK) {
        std::remove(tmp.c_str());
        return rc;
    }
    if (std::rename(tmp.c_str(), path) != 0) return fail(std::string("cannot rename to ") + path);
    return NEMO_FLASH_OK;
}

int nemo_flash_error_message(char* buf, int32_t len) {
    if (!buf || len <= 0) return NEMO_FLASH_ERR;
    std::snprintf(buf, size_t(len), "%s", last_error().c_str());
    return NEMO_FLASH_OK;
}

}  // extern "C"
