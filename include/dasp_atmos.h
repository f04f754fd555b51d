/*
 * dasp_atmos.h - C ABI of the device phase producer upstream of the Shack-Hartmann path: the
 * atmosphere's pupil phase and the deformable mirror's surface on the line of sight, kept in HBM
 * so that phase never leaves the device between the atmosphere, the mirror and the WFS engine
 * (include/dasp_cent.h). Library: dasp_amd/libdasp_cent.so. All pointers named d_* are device
 * memory; every call is asynchronous on the given HIP stream (NULL: the null stream). Return
 * values: 0 or a negative DASP_CENT_E* code (dasp_cent.h), with the message in
 * dasp_atmos_last_error().
 *
 * Reference counterparts (agb32/dasp): util/atmos.py atmos.createPupilPhs (:678-833), infAtmos
 * makeLayer (science/infAtmos.py:364-402), cmod/src/interpmodule.c interp_linearshift
 * (:1234-1300) and gslCubSplineIntrp (:294-600), util/dm.py MirrorSurface (:1268-1740) and
 * DMLineOfSight.selectSubPupil (:2057-2151), science/xinterp_dm.py generateNext (:407-513).
 */
#ifndef DASP_ATMOS_H
#define DASP_ATMOS_H

#ifdef __cplusplus
extern "C" {
#endif

#define DASP_ATMOS_MAX_LAYERS 16

/* Message of the calling thread's last failed dasp_atmos_* / dasp_dm_* call. */
const char *dasp_atmos_last_error(void);

/* One atmosphere layer as createPupilPhs reads it for this step. The screen is a ring buffer: the
 * logical element (r, c) of the reference's phaseScreens[key] sits at physical
 * ((r + oy) mod rows, (c + ox) mod cols), so that makeLayer's shift of the whole screen becomes a
 * move of the origin (dasp_atmos_layer_update). r0, c0: the logical window start
 * (posDict[1] + phsShiftX, posDict[0] + phsShiftY, util/atmos.py:697-725); s1..s4: linearshift's
 * float weights for this step's fractional shifts (interp_linearshift :1276-1279). */
typedef struct {
  const double *screen;
  int rows, cols;
  int oy, ox;
  int r0, c0;
  float s1, s2, s3, s4;
} dasp_atmos_layer;

/* Doubles of device work space dasp_atmos_pupil_phase needs for an npup x npup pupil. */
long dasp_atmos_work_doubles(int npup);

/* createPupilPhs for a natural guide star (util/atmos.py:678-833 with sourceAlt 0, no zenith
 * rebinning, no uplink tip-tilt): out = intrinsic (or 0) + the layers' linearshift windows added in
 * the given order (float64); then, if remove_piston, out -= sum(out * pfn) / pupil_area
 * (control["removePiston"]); if d_pupfn is non-NULL, out *= pfn (d_pupfn NULL stands for
 * control["fullPupil"], pupil_area then npup^2); out *= out_scale (500/sourceLam, 1 to skip).
 * The result goes to d_out64 (float64, the reference's outputData) and/or d_out32 (its float32
 * cast, what a WFS copies in, science/wfscent.py:680-682); either may be NULL, not both. */
int dasp_atmos_pupil_phase(const dasp_atmos_layer *layers, int nlayers, int npup, const double *d_intrinsic,
                           const float *d_pupfn, double pupil_area, int remove_piston, double out_scale,
                           double *d_out64, float *d_out32, double *d_work, void *stream);

/* makeLayer (science/infAtmos.py:364-402) on a ring-buffered screen: the screen moves by nrows
 * rows and ncols columns away from the side new data enters (col_add < 0: new columns at the end,
 * else at the start; likewise row_add), the new columns d_cols (ncols x rows, colInput's layout)
 * are written, then the new rows d_rows (nrows x cols) over the full width. *oy, *ox: the origin,
 * updated in place. */
int dasp_atmos_layer_update(double *d_screen, int rows, int cols, int *oy, int *ox, const double *d_cols,
                            int ncols, const double *d_rows, int nrows, int col_add, int row_add, void *stream);

/* Host: the natural cubic spline operator of MirrorSurface's "spline" type (util/dm.py:1467-1486,
 * :1520-1536): op[k * nact + a] = the value at pixel x_k = k of the natural spline through the
 * actuator positions x2_a = a * (npup - 1) / (nact - 1 + 2 actoffset) + actoffset * spacing with
 * value 1 at actuator a and 0 elsewhere (GSL's gsl_interp_cspline as gslCubSplineIntrp uses it).
 * Pixels outside the actuator span (actoffset > 0) are refused, as GSL refuses them. op: host
 * (npup, nact) doubles. */
int dasp_dm_spline_operator(int npup, int nact, double actoffset, double *op);

/* MirrorSurface.fit, "spline" type: the actuator map d_act (nact x nact, float32) coupled to its
 * neighbours (MirrorSurface.fudge, util/dm.py:1723-1740, float32 arithmetic; coupling 0 skips
 * it), then interpolated as gslCubSplineIntrp does, in double: surface = B F B^T with B the
 * operator above (d_op (npup, nact) and its transpose d_opT (nact, npup), float64), cast to float32
 * into d_surface (npup x npup). d_work: nact * nact floats + npup * nact doubles. */
int dasp_dm_surface(const float *d_act, int nact, float coupling, const double *d_op, const double *d_opT, int npup,
                    float *d_surface, void *d_work, void *stream);

/* DMLineOfSight.selectSubPupil for a ground-conjugate mirror without sub-pixel interpolation
 * (util/dm.py:2057-2151), float32 as xinterp_dm's outputData: out = base + factor * surface[yoff + i,
 * xoff + j] (factor = wavelengthAdjustor, negated for conjHeight < 0; the product is formed only
 * when the adjustor is not 1), base = the float32 cast of d_base64, or d_base32, or 0 when both
 * are NULL (no atmosphere parent); then out *= pfn, and with remove_piston out -= sum(out) /
 * pupil_sum (the sum accumulated in double), out *= pfn. d_work: dasp_atmos_work_doubles(npup). */
int dasp_dm_line_of_sight(const float *d_surface, int dmpup, int yoff, int xoff, int npup, float factor,
                          int adjust, const double *d_base64, const float *d_base32, const float *d_pupfn,
                          double pupil_sum, int remove_piston, float *d_out32, double *d_work, void *stream);

#ifdef __cplusplus
}
#endif
#endif
