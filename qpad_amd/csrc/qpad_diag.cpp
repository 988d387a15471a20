// qpad_diag.cpp -- QPAD's HDF5 diagnostics from libqpad_hip's device state.
//
// Host code on the HDF5 1.10 C API.  The data come through libqpad_hip's C
// ABI (histories, species and beam arrays); the file layout follows QPAD's
// writer: hdf5io_class.f03:264-335 (file attributes, AXIS group),
// :420-470 (2D grid dataset), :1584-1700 (plasma raw), :1220-1582 (beam
// raw); file names and the per-mode folders follow diagnostics_class.f03
// (init_diag_fields :396-606, species and beams :180-330, add_diag_cym
// :947-1043).
#include <errno.h>
#include <hdf5.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/qpad_diag.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

// QPAD's hdf5file string fields are character(len=100): attributes written
// from them are 100 characters, space padded
constexpr size_t kFortranLen = 100;

// mkdir -p of the folder part of `path`
int make_dirs(const std::string &path)
{
    for (size_t p = 1; p < path.size(); p++) {
        if (path[p] != '/') continue;
        const std::string d = path.substr(0, p);
        if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) return fail(QPAD_E_INVALID, "cannot create " + d);
    }
    return 0;
}

// a fixed-length, space-padded string attribute of `len` characters
// (add_h5_atribute_str: H5T_NATIVE_CHARACTER of the Fortran string's length)
bool attr_str(hid_t obj, const char *name, const std::string &value, size_t len)
{
    std::string v = value.substr(0, len);
    v.resize(len, ' ');
    const hid_t t = H5Tcopy(H5T_FORTRAN_S1);
    H5Tset_size(t, len);
    const hsize_t one = 1;
    const hid_t sp = H5Screate_simple(1, &one, nullptr);
    const hid_t a = H5Acreate2(obj, name, t, sp, H5P_DEFAULT, H5P_DEFAULT);
    const bool ok = a >= 0 && H5Awrite(a, t, v.data()) >= 0;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    H5Tclose(t);
    return ok;
}

bool attr_double(hid_t obj, const char *name, double value)
{
    const hsize_t one = 1;
    const hid_t sp = H5Screate_simple(1, &one, nullptr);
    const hid_t a = H5Acreate2(obj, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
    const bool ok = a >= 0 && H5Awrite(a, H5T_NATIVE_DOUBLE, &value) >= 0;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return ok;
}

bool attr_int(hid_t obj, const char *name, int value)
{
    const hsize_t one = 1;
    const hid_t sp = H5Screate_simple(1, &one, nullptr);
    const hid_t a = H5Acreate2(obj, name, H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT);
    const bool ok = a >= 0 && H5Awrite(a, H5T_NATIVE_INT, &value) >= 0;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return ok;
}

// what QPAD's hdf5file holds for one file
struct FileAttrs {
    std::string dataname, type, units, label;
    int iter = 0;
    double dt = 1.0;
    bool grid = true;
    double axmin[2] = {0.0, 0.0}, axmax[2] = {1.0, 1.0};
};

// wrattr_file (hdf5io_class.f03:264-313)
bool write_file_attrs(hid_t file, const FileAttrs &f)
{
    const hid_t root = H5Gopen2(file, "/", H5P_DEFAULT);
    bool ok = root >= 0;
    ok = ok && attr_str(root, "NAME", f.dataname, kFortranLen);
    ok = ok && attr_str(root, "TYPE", f.type, kFortranLen);
    ok = ok && attr_double(root, "TIME", f.iter * f.dt);
    ok = ok && attr_int(root, "ITER", f.iter);
    ok = ok && attr_double(root, "DT", f.dt);
    ok = ok && attr_str(root, "TIME UNIT", "1 / \\omega_p", kFortranLen);
    if (ok && f.grid) {
        const hid_t ax = H5Gcreate2(root, "AXIS", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        ok = ax >= 0;
        const char *names[2] = {"r", "\\xi"};
        for (int i = 0; ok && i < 2; i++) {
            const hsize_t two = 2;
            const hid_t sp = H5Screate_simple(1, &two, nullptr);
            const std::string dn = "AXIS" + std::to_string(i + 1);
            const hid_t d = H5Dcreate2(ax, dn.c_str(), H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
            ok = d >= 0;
            const std::string nm = names[i];
            ok = ok && attr_str(d, "TYPE", "linear", 6);
            ok = ok && attr_str(d, "UNITS", "c / \\omega_p", 12);
            ok = ok && attr_str(d, "NAME", nm, nm.size());
            ok = ok && attr_str(d, "LONG_NAME", nm, nm.size());
            const double v[2] = {f.axmin[i], f.axmax[i]};
            ok = ok && H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) >= 0;
            if (d >= 0) H5Dclose(d);
            H5Sclose(sp);
        }
        if (ax >= 0) H5Gclose(ax);
    }
    if (root >= 0) H5Gclose(root);
    return ok;
}

// one double dataset of rank 1 or 2 with QPAD's UNITS / LONG_NAME
// attributes (wrattr_dataset, hdf5io_class.f03:315-335); ulen / llen: the
// attribute lengths (100 for the hdf5file fields, the literal's for the
// particle datasets)
bool write_dataset(hid_t loc, const char *name, const double *data, int rank, const hsize_t *dims,
                   const std::string &units, size_t ulen, const std::string &label, size_t llen)
{
    const hid_t sp = H5Screate_simple(rank, dims, nullptr);
    const hid_t d = H5Dcreate2(loc, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    bool ok = d >= 0 && H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
    ok = ok && attr_str(d, "UNITS", units, ulen);
    ok = ok && attr_str(d, "LONG_NAME", label, llen);
    if (d >= 0) H5Dclose(d);
    H5Sclose(sp);
    return ok;
}

std::string iter_file(const std::string &folder, const std::string &dataname, int iter)
{
    char st[16];
    snprintf(st, sizeof st, "%08d", iter);
    return folder + dataname + "_" + st + ".h5";
}

// the per-mode folder of mode array k (add_diag_cym, dtype 'real'): Re0,
// then Re m (k = 2m - 1) and Im m (k = 2m)
std::string mode_folder(int k)
{
    if (k == 0) return "Re0";
    return ((k & 1) ? "Re" : "Im") + std::to_string((k + 1) / 2);
}

int write_grid_file(const std::string &path, const FileAttrs &fa, const double *data, int nxi, int nr)
{
    if (make_dirs(path)) return QPAD_E_INVALID;
    const hid_t file = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) return fail(QPAD_E_INVALID, "cannot create " + path);
    bool ok = write_file_attrs(file, fa);
    const hid_t root = H5Gopen2(file, "/", H5P_DEFAULT);
    const hsize_t dims[2] = {(hsize_t)nxi, (hsize_t)nr};
    ok = ok && root >= 0 && write_dataset(root, fa.dataname.c_str(), data, 2, dims, fa.units, kFortranLen, fa.label,
                                          kFortranLen);
    if (root >= 0) H5Gclose(root);
    ok = H5Fclose(file) >= 0 && ok;
    return ok ? 0 : fail(QPAD_E_INVALID, "HDF5 write failed for " + path);
}

// the field history `fid` component `comp` of every slice into one
// (nxi, nr) array per mode array; get(s, buf) fills slice s as
// [ncomp][nm][nr + 2]
template <class Get>
int write_mode_files(qpad_ctx *c, Get get, int ncomp, int comp, const std::string &folder,
                     const std::string &dataname, const std::string &units, const std::string &label, int iter,
                     const qpad_diag_box *box, int nr, int nxi, int nm)
{
    const size_t w = (size_t)nr + 2;
    std::vector<double> slice((size_t)ncomp * nm * w), out((size_t)nm * nxi * nr);
    for (int s = 1; s <= nxi; s++) {
        const int rc = get(s, slice.data());
        if (rc) return fail(rc, std::string("reading the history: ") + qpad_last_error());
        for (int k = 0; k < nm; k++) {
            const double *row = slice.data() + ((size_t)comp * nm + k) * w;
            memcpy(out.data() + ((size_t)k * nxi + (s - 1)) * nr, row + 1, sizeof(double) * nr);
        }
    }
    (void)c;
    FileAttrs fa;
    fa.dataname = dataname;
    fa.type = "grid";
    fa.units = units;
    fa.label = label;
    fa.iter = iter;
    fa.dt = box->dt;
    fa.axmin[0] = box->rmin;
    fa.axmax[0] = box->rmax;
    fa.axmin[1] = box->zmin;
    fa.axmax[1] = box->zmax;
    for (int k = 0; k < nm; k++) {
        const std::string path = iter_file(folder + mode_folder(k) + "/", dataname, iter);
        const int rc = write_grid_file(path, fa, out.data() + (size_t)k * nxi * nr, nxi, nr);
        if (rc) return rc;
    }
    return 0;
}

struct FieldDiag {
    const char *name, *folder, *dataname, *units, *label;
    int fid, ncomp, comp;
};

// init_diag_fields (diagnostics_class.f03:424-580)
const FieldDiag kFieldDiags[] = {
    {"er_cyl_m", "Er", "er", "m_ec\\omega_p/e", "E_r", QPAD_F_E, 3, 0},
    {"ephi_cyl_m", "Ephi", "ephi", "m_ec\\omega_p/e", "E_\\phi", QPAD_F_E, 3, 1},
    {"ez_cyl_m", "Ez", "ez", "m_ec\\omega_p/e", "E_z", QPAD_F_E, 3, 2},
    {"br_cyl_m", "Br", "br", "m_e\\omega_p/e", "B_r", QPAD_F_B, 3, 0},
    {"bphi_cyl_m", "Bphi", "bphi", "m_e\\omega_p/e", "B_\\phi", QPAD_F_B, 3, 1},
    {"bz_cyl_m", "Bz", "bz", "m_e\\omega_p/e", "B_z", QPAD_F_B, 3, 2},
    {"spec_er_cyl_m", "Er_spec", "er", "m_ec\\omega_p/e", "E_r", QPAD_F_E_SPE, 3, 0},
    {"spec_ephi_cyl_m", "Ephi_spec", "ephi", "m_ec\\omega_p/e", "E_\\phi", QPAD_F_E_SPE, 3, 1},
    {"spec_br_cyl_m", "Br_spec", "br", "m_e\\omega_p/e", "B_r", QPAD_F_B_SPE, 3, 0},
    {"spec_bphi_cyl_m", "Bphi_spec", "bphi", "m_e\\omega_p/e", "B_\\phi", QPAD_F_B_SPE, 3, 1},
    {"spec_bz_cyl_m", "Bz_spec", "bz", "m_e\\omega_p/e", "B_z", QPAD_F_B_SPE, 3, 2},
    {"charge_cyl_m", "Charge", "charge", "n_0", "\\rho", QPAD_F_Q_SPE, 1, 0},
    {"jr_cyl_m", "Jr", "jr", "n_0 c", "J_r", QPAD_F_CU, 3, 0},
    {"jphi_cyl_m", "Jphi", "jphi", "n_0 c", "J_\\phi", QPAD_F_CU, 3, 1},
    {"jz_cyl_m", "Jz", "jz", "n_0 c", "J_z", QPAD_F_CU, 3, 2},
    {"psi_cyl_m", "Psi", "psi", "m_ec^2/e", "\\Psi", QPAD_F_PSI, 1, 0},
};

// grid sizes of a context (qpad_get_grid)
struct Shape {
    int nr, nxi, nm;
};

bool valid_box(const qpad_diag_box *b) { return b && b->rmax > b->rmin && b->zmax > b->zmin; }

std::string slash(const char *dir) { return (dir && *dir) ? std::string(dir) + "/" : std::string("./"); }

}  // namespace

extern "C" {

const char *qpad_diag_last_error(void) { return g_err.c_str(); }

int qpad_diag_write_grid(const char *path, const char *dataname, const double *data, int32_t nxi, int32_t nr,
                         int32_t iter, const qpad_diag_box *box, const char *units, const char *label)
{
    if (!path || !dataname || !data || nxi < 1 || nr < 1 || !valid_box(box))
        return fail(QPAD_E_INVALID, "bad grid-file arguments");
    FileAttrs fa;
    fa.dataname = dataname;
    fa.type = "grid";
    fa.units = units ? units : "";
    fa.label = label ? label : "";
    fa.iter = iter;
    fa.dt = box->dt;
    fa.axmin[0] = box->rmin;
    fa.axmax[0] = box->rmax;
    fa.axmin[1] = box->zmin;
    fa.axmax[1] = box->zmax;
    return write_grid_file(path, fa, data, nxi, nr);
}

}  // extern "C"

namespace {

int grid_of(qpad_ctx *c, Shape *sh)
{
    qpad_grid g;
    const int rc = qpad_get_grid(c, &g);
    if (rc) return fail(rc, std::string("qpad_get_grid: ") + qpad_last_error());
    sh->nr = g.nr;
    sh->nxi = g.nxi;
    sh->nm = 2 * g.max_mode + 1;
    return 0;
}

}  // namespace

extern "C" {

int qpad_diag_field(qpad_ctx *c, const char *name, const char *dir, int32_t iter, const qpad_diag_box *box)
{
    if (!c || !name || !valid_box(box)) return fail(QPAD_E_INVALID, "bad field-diagnostic arguments");
    const FieldDiag *fd = nullptr;
    for (const FieldDiag &d : kFieldDiags)
        if (!strcmp(d.name, name)) fd = &d;
    if (!fd) return fail(QPAD_E_INVALID, std::string("field diagnostic '") + name + "' is not on the GPU path");
    Shape sh;
    if (grid_of(c, &sh)) return QPAD_E_INVALID;
    auto get = [&](int s, double *buf) { return qpad_get_hist(c, fd->fid, s, buf); };
    return write_mode_files(c, get, fd->ncomp, fd->comp, slash(dir) + "Fields/" + fd->folder + "/", fd->dataname,
                            fd->units, fd->label, iter, box, sh.nr, sh.nxi, sh.nm);
}

// the charge history of species sp under <dir>/<folder>/Charge/
static int species_charge_files(qpad_ctx *c, int32_t sp, const std::string &folder, const char *dir, int32_t iter,
                                const qpad_diag_box *box)
{
    Shape sh;
    if (grid_of(c, &sh)) return QPAD_E_INVALID;
    auto get = [&](int s, double *buf) { return qpad_get_species_charge(c, sp, s, buf); };
    return write_mode_files(c, get, 1, 0, slash(dir) + folder + "/Charge/", "charge", "n_0", "\\rho", iter, box,
                            sh.nr, sh.nxi, sh.nm);
}

int qpad_diag_species_charge(qpad_ctx *c, int32_t sp, const char *dir, int32_t iter, const qpad_diag_box *box)
{
    if (!c || sp < 0 || !valid_box(box)) return fail(QPAD_E_INVALID, "bad species-charge arguments");
    return species_charge_files(c, sp, "Species" + std::to_string(sp + 1), dir, iter, box);
}

int qpad_diag_neutral_charge(qpad_ctx *c, int32_t sp, int32_t nt, const char *dir, int32_t iter,
                             const qpad_diag_box *box)
{
    if (!c || sp < 0 || nt < 0 || !valid_box(box)) return fail(QPAD_E_INVALID, "bad neutral-charge arguments");
    return species_charge_files(c, sp, "Neutral" + std::to_string(nt + 1), dir, iter, box);
}

int qpad_diag_neutral_ion(qpad_ctx *c, int32_t sp, int32_t nt, const char *dir, int32_t iter,
                          const qpad_diag_box *box)
{
    if (!c || sp < 0 || nt < 0 || !valid_box(box)) return fail(QPAD_E_INVALID, "bad neutral-ion arguments");
    Shape sh;
    if (grid_of(c, &sh)) return QPAD_E_INVALID;
    auto get = [&](int s, double *buf) { return qpad_get_ion_density(c, sp, s, buf); };
    return write_mode_files(c, get, 1, 0, slash(dir) + "Neutral" + std::to_string(nt + 1) + "/Ion_charge/",
                            "ion_charge", "n_0", "\\rho", iter, box, sh.nr, sh.nxi, sh.nm);
}

int qpad_diag_beam_charge(qpad_ctx *c, int32_t ib, const char *dir, int32_t iter, const qpad_diag_box *box)
{
    if (!c || ib < 0 || !valid_box(box)) return fail(QPAD_E_INVALID, "bad beam-charge arguments");
    int64_t n = 0;
    if (qpad_beam_count(c, ib, &n)) return fail(QPAD_E_INVALID, std::string("beam: ") + qpad_last_error());
    // q_beam is the sum over the beams: it is beam ib's own only when the
    // context holds that one beam
    if (ib != 0 || qpad_beam_count(c, 1, &n) == 0)
        return fail(QPAD_E_INVALID, "beam charge_cyl_m needs a context with one beam (q_beam sums every beam)");
    Shape sh;
    if (grid_of(c, &sh)) return QPAD_E_INVALID;
    auto get = [&](int s, double *buf) { return qpad_get_hist(c, QPAD_F_Q_BEAM, s, buf); };
    return write_mode_files(c, get, 1, 0, slash(dir) + "Beam" + std::to_string(ib + 1) + "/Charge/", "charge",
                            "n_0", "\\rho", iter, box, sh.nr, sh.nxi, sh.nm);
}

// every psample-th particle of species sp into <dir>/<folder>/Raw/
static int species_raw_file(qpad_ctx *c, int32_t sp, int32_t psample, const std::string &folder, const char *label,
                            const char *dir, int32_t iter, const qpad_diag_box *box);

int qpad_diag_species_raw(qpad_ctx *c, int32_t sp, int32_t psample, const char *dir, int32_t iter,
                          const qpad_diag_box *box)
{
    if (!c || sp < 0 || psample < 1 || !valid_box(box)) return fail(QPAD_E_INVALID, "bad species-raw arguments");
    return species_raw_file(c, sp, psample, "Species" + std::to_string(sp + 1), "Species Raw", dir, iter, box);
}

int qpad_diag_neutral_raw(qpad_ctx *c, int32_t sp, int32_t nt, int32_t psample, const char *dir, int32_t iter,
                          const qpad_diag_box *box)
{
    if (!c || sp < 0 || nt < 0 || psample < 1 || !valid_box(box))
        return fail(QPAD_E_INVALID, "bad neutral-raw arguments");
    return species_raw_file(c, sp, psample, "Neutral" + std::to_string(nt + 1), "Neutral Raw", dir, iter, box);
}

static int species_raw_file(qpad_ctx *c, int32_t sp, int32_t psample, const std::string &folder, const char *label,
                            const char *dir, int32_t iter, const qpad_diag_box *box)
{
    int64_t n = 0;
    if (qpad_species_count(c, sp, &n)) return fail(QPAD_E_INVALID, std::string("species: ") + qpad_last_error());
    std::vector<double> a[8];
    for (auto &v : a) v.resize((size_t)std::max<int64_t>(n, 1));
    std::vector<int32_t> id((size_t)std::max<int64_t>(n, 1));
    if (qpad_get_species(c, sp, a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data(), a[5].data(),
                         a[6].data(), a[7].data(), id.data()))
        return fail(QPAD_E_INVALID, std::string("species: ") + qpad_last_error());
    // wpart (hdf5io_class.f03:1584-1700): tp = npp / dspl, rows of the
    // Fortran (2, tp) / (3, tp) arrays are particles
    const int64_t tp = n / psample;
    const std::string path = iter_file(slash(dir) + folder + "/Raw/", "raw", iter);
    if (make_dirs(path)) return QPAD_E_INVALID;
    const hid_t file = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) return fail(QPAD_E_INVALID, "cannot create " + path);
    FileAttrs fa;
    fa.dataname = "raw";
    fa.type = "particles";
    fa.iter = iter;
    fa.dt = box->dt;
    fa.grid = false;
    bool ok = write_file_attrs(file, fa);
    const hid_t root = H5Gopen2(file, "/", H5P_DEFAULT);
    ok = ok && root >= 0 && attr_int(root, "tp", (int)tp);
    if (ok && tp > 0) {
        std::vector<double> x((size_t)tp * 2), p((size_t)tp * 3), q((size_t)tp);
        for (int64_t t = 0; t < tp; t++) {
            const size_t i = (size_t)(t * psample);
            x[2 * t] = a[0][i];
            x[2 * t + 1] = a[1][i];
            p[3 * t] = a[2][i];
            p[3 * t + 1] = a[3][i];
            p[3 * t + 2] = a[4][i];
            q[t] = a[5][i];
        }
        const hsize_t dx[2] = {(hsize_t)tp, 2}, dp[2] = {(hsize_t)tp, 3}, dq[1] = {(hsize_t)tp};
        ok = ok && write_dataset(root, "x", x.data(), 2, dx, "", kFortranLen, label, kFortranLen);
        ok = ok && write_dataset(root, "p", p.data(), 2, dp, "", kFortranLen, label, kFortranLen);
        ok = ok && write_dataset(root, "q", q.data(), 1, dq, "", kFortranLen, label, kFortranLen);
    }
    if (root >= 0) H5Gclose(root);
    ok = H5Fclose(file) >= 0 && ok;
    return ok ? 0 : fail(QPAD_E_INVALID, "HDF5 write failed for " + path);
}

int qpad_diag_beam_raw(qpad_ctx *c, int32_t ib, int32_t psample, const char *dir, int32_t iter,
                       const qpad_diag_box *box)
{
    if (!c || ib < 0 || psample < 1 || !valid_box(box)) return fail(QPAD_E_INVALID, "bad beam-raw arguments");
    int64_t n = 0;
    if (qpad_beam_count(c, ib, &n)) return fail(QPAD_E_INVALID, std::string("beam: ") + qpad_last_error());
    std::vector<double> a[7];
    for (auto &v : a) v.resize((size_t)std::max<int64_t>(n, 1));
    if (qpad_get_beam(c, ib, a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data(), a[5].data(),
                      a[6].data()))
        return fail(QPAD_E_INVALID, std::string("beam: ") + qpad_last_error());
    // pwpart_3d_pipe (hdf5io_class.f03:1220-1582): x1, x2, x3 + z0, p1..p3, q
    const int64_t tp = n / psample;
    const std::string path = iter_file(slash(dir) + "Beam" + std::to_string(ib + 1) + "/Raw/", "raw", iter);
    if (make_dirs(path)) return QPAD_E_INVALID;
    const hid_t file = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) return fail(QPAD_E_INVALID, "cannot create " + path);
    FileAttrs fa;
    fa.dataname = "raw";
    fa.type = "particles";
    fa.iter = iter;
    fa.dt = box->dt;
    fa.grid = false;
    bool ok = write_file_attrs(file, fa);
    const hid_t root = H5Gopen2(file, "/", H5P_DEFAULT);
    ok = ok && root >= 0 && attr_int(root, "tp", (int)tp);
    if (ok && tp > 0) {
        std::vector<double> buf((size_t)tp);
        const char *names[7] = {"x1", "x2", "x3", "p1", "p2", "p3", "q"};
        const char *labels[7] = {"x_1", "x_2", "x_3", "p_1", "p_2", "p_3", "q"};
        const char *units[7] = {"c/\\omega_p", "c/\\omega_p", "c/\\omega_p", "m_ec", "m_ec", "m_ec", "a.u."};
        const hsize_t d1[1] = {(hsize_t)tp};
        for (int f = 0; ok && f < 7; f++) {
            for (int64_t t = 0; t < tp; t++) buf[t] = a[f][(size_t)(t * psample)] + (f == 2 ? box->zmin : 0.0);
            const std::string u = units[f], l = labels[f];
            ok = write_dataset(root, names[f], buf.data(), 1, d1, u, u.size(), l, l.size());
        }
    }
    if (root >= 0) H5Gclose(root);
    ok = H5Fclose(file) >= 0 && ok;
    return ok ? 0 : fail(QPAD_E_INVALID, "HDF5 write failed for " + path);
}

int qpad_diag_read_beam_file(const char *path, int64_t *n, double *x1, double *x2, double *x3, double *p1,
                             double *p2, double *p3, double *q)
{
    if (!path || !n) return fail(QPAD_E_INVALID, "bad beam-file arguments");
    const hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return fail(QPAD_E_INVALID, std::string("cannot open ") + path);
    const char *names[7] = {"x1", "x2", "x3", "p1", "p2", "p3", "q"};
    double *outs[7] = {x1, x2, x3, p1, p2, p3, q};
    int rc = 0;
    int64_t count = -1;
    for (int f = 0; f < 7 && !rc; f++) {
        const hid_t d = H5Dopen2(file, names[f], H5P_DEFAULT);
        if (d < 0) {
            rc = fail(QPAD_E_INVALID, std::string("no dataset ") + names[f] + " in " + path);
            break;
        }
        const hid_t sp = H5Dget_space(d);
        const int rank = H5Sget_simple_extent_ndims(sp);
        hsize_t dims[2] = {0, 0};
        int64_t len = -1;
        if (rank == 1 || rank == 2) {
            H5Sget_simple_extent_dims(sp, dims, nullptr);
            if (rank == 1) len = (int64_t)dims[0];
            else if (dims[0] == 1 || dims[1] == 1) len = (int64_t)(dims[0] * dims[1]);
        }
        if (len < 0) rc = fail(QPAD_E_INVALID, std::string("dataset ") + names[f] + " is not a list of values");
        else if (count >= 0 && len != count) rc = fail(QPAD_E_INVALID, "the beam datasets differ in length");
        count = len;
        if (!rc && x1) {
            if (len != *n) rc = fail(QPAD_E_INVALID, "array length differs from the file's particle count");
            else if (len > 0 && H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, outs[f]) < 0)
                rc = fail(QPAD_E_INVALID, std::string("reading ") + names[f] + " failed");
        }
        H5Sclose(sp);
        H5Dclose(d);
    }
    H5Fclose(file);
    if (!rc && !x1) *n = count;
    return rc;
}

}  // extern "C"
