/*
 * qpad_diag.h -- QPAD's HDF5 diagnostics written from libqpad_hip's device
 * state (libqpad_diag.so, built on HDF5 1.10 and the C ABI of qpad_hip.h).
 *
 * The files reproduce the layout of QPAD's own writer so that its
 * post-processing (tools/merge_modes_tools) reads them unchanged:
 *   - field and charge diagnostics *_cyl_m (diagnostics_class.f03:396-606,
 *     add_diag_cym :947-1043): one file per mode array, in
 *     <dir>/<Fields/Name | SpeciesN/Charge | BeamN/Charge>/<Re0|Re1|Im1|..>/
 *     <dataname>_<iter, 8 digits>.h5, each a 2D dataset <dataname> of shape
 *     (nxi, nr) (QPAD's Fortran (nr, nxi)), nodes 1..nr of slices 1..nxi;
 *     root attributes NAME, TYPE, TIME, ITER, DT, TIME UNIT, group AXIS with
 *     AXIS1 = (rmin, rmax), AXIS2 = (zmin, zmax) (hdf5io_class.f03:264-313,
 *     420-470);
 *   - raw particles (hdf5io_class.f03:1584-1700 for plasma species,
 *     :1220-1582 for beams): attribute tp and the particle datasets.
 * String attributes are space-padded fixed-length strings of QPAD's lengths
 * (its hdf5file fields are character(len=100)).
 *
 * Every call returns 0 or a negative code (qpad_hip.h); the message is in
 * qpad_diag_last_error().  No call aborts or prints.
 */
#ifndef QPAD_DIAG_H
#define QPAD_DIAG_H

#include <stdint.h>

#include "qpad_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* simulation.box and simulation.dt of the deck (the axis ranges and time of
   every file) */
typedef struct {
    double rmin, rmax, zmin, zmax;
    double dt;
} qpad_diag_box;

const char *qpad_diag_last_error(void);

/* one (nxi, nr) grid file at `path` (directories created): data[j][i],
   j = slice 0..nxi-1, i = node 0..nr-1; attributes as QPAD's hdf5file
   (iter = ITER, iter * dt = TIME) */
int qpad_diag_write_grid(const char *path, const char *dataname, const double *data, int32_t nxi, int32_t nr,
                         int32_t iter, const qpad_diag_box *box, const char *units, const char *label);

/* field diagnostic `name` of the 3D step just completed (QPAD's
   field.diag names: er_cyl_m, ephi_cyl_m, ez_cyl_m, br_cyl_m, bphi_cyl_m,
   bz_cyl_m, spec_er/ephi/br/bphi/bz_cyl_m, charge_cyl_m, jr/jphi/jz_cyl_m,
   psi_cyl_m) from the slice histories: <dir>/Fields/<Name>/... */
int qpad_diag_field(qpad_ctx *ctx, const char *name, const char *dir, int32_t iter, const qpad_diag_box *box);

/* species(sp+1) charge_cyl_m: <dir>/Species<sp+1>/Charge/... from the
   per-species charge history (qpad_species_charge_history must be on) */
int qpad_diag_species_charge(qpad_ctx *ctx, int32_t sp, const char *dir, int32_t iter, const qpad_diag_box *box);

/* beam(ib+1) charge_cyl_m: <dir>/Beam<ib+1>/Charge/... (contexts with one
   beam: the q_beam history is that beam's) */
int qpad_diag_beam_charge(qpad_ctx *ctx, int32_t ib, const char *dir, int32_t iter, const qpad_diag_box *box);

/* raw particles, every psample-th one: species(sp+1) into
   <dir>/Species<sp+1>/Raw/raw_<iter>.h5 (x (tp, 2), p (tp, 3), q (tp)),
   beam(ib+1) into <dir>/Beam<ib+1>/Raw/raw_<iter>.h5 (x1, x2, x3 + zmin, p1,
   p2, p3, q).  The library keeps particles sorted by cell, so the sample is
   every psample-th particle in that order. */
int qpad_diag_species_raw(qpad_ctx *ctx, int32_t sp, int32_t psample, const char *dir, int32_t iter,
                          const qpad_diag_box *box);
int qpad_diag_beam_raw(qpad_ctx *ctx, int32_t ib, int32_t psample, const char *dir, int32_t iter,
                       const qpad_diag_box *box);

/* neutrals(nt+1) diagnostics (diagnostics_class.f03:324-390, 909-928) of
   the neutral whose species id is sp: charge_cyl_m (its electrons' charge
   history, qpad_species_charge_history on) into <dir>/Neutral<nt+1>/Charge/,
   ion_cyl_m (rho_ion, neutral%wr_ion) into <dir>/Neutral<nt+1>/Ion_charge/
   (dataname ion_charge), raw (its electrons) into <dir>/Neutral<nt+1>/Raw/ */
int qpad_diag_neutral_charge(qpad_ctx *ctx, int32_t sp, int32_t nt, const char *dir, int32_t iter,
                             const qpad_diag_box *box);
int qpad_diag_neutral_ion(qpad_ctx *ctx, int32_t sp, int32_t nt, const char *dir, int32_t iter,
                          const qpad_diag_box *box);
int qpad_diag_neutral_raw(qpad_ctx *ctx, int32_t sp, int32_t nt, int32_t psample, const char *dir, int32_t iter,
                          const qpad_diag_box *box);

/* the particle file of a "file" beam (fdist3d_file_class.f03:154-343):
   datasets x1, x2, x3, p1, p2, p3, q of the root group, each holding n
   values (rank 1, or rank 2 with one extent 1).  With x1 null only n is
   returned; otherwise every array must hold *n values. */
int qpad_diag_read_beam_file(const char *path, int64_t *n, double *x1, double *x2, double *x3, double *p1,
                             double *p2, double *p3, double *q);

#ifdef __cplusplus
}
#endif
#endif
