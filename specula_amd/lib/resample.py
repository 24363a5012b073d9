"""Field resampling of SH for fov_ovs != 1, rotation or shift (host-side set-up).

SH.setup (specula/processing_objects/sh.py:448-458) builds, once per pupil, the edge-
extrapolation maps of extrapolate_edge_pixel_mat_define (lib/extrapolate_edge_pixel_mat_define.py:
5-140) and the sampling points of Interp2D (lib/interp2d.py:48-93); SH.trigger_code
(sh.py:350-360) then extrapolates the phase 1-2 pixels beyond the pupil edge and resamples A and
the extrapolated phase bilinearly on the ef*fov_ovs grid.  The per-frame part runs in
``resample_kernel`` (csrc/shwfs.hip); this module prepares its inputs.
"""
import math

import numpy as np

# direction order of the reference (its "y+", "y-", "x+", "x-": columns first, then rows)
_DIRS = ((0, 1), (0, -1), (1, 0), (-1, 0))


def edge_extrapolation_maps(mask, two_pixel=True):
    """Source pixels (linear indices, -1 = none) of the 1-pixel and 2-pixel linear
    extrapolation of every pixel just outside ``mask`` (the pupil amplitude).  A pixel p with
    src1[p] >= 0 takes phase[src1[p]], or 2 phase[src1[p]] - phase[src2[p]] when src2[p] >= 0.

    Same selection rules as the reference (periodic neighbour test, strict border limits, the
    four directions tried in order, a single-pixel source kept only when no direction offers
    two), evaluated for all pixels at once.  As there, the amplitude values themselves weigh the
    neighbour test ((sum of neighbours) * (1 - A) > 0): a pixel with 0 < A < 1 next to the pupil
    is extrapolated too; a 0/1 pupil reduces to the plain edge ring."""
    f = np.asarray(mask, dtype=np.float64)
    H, W = f.shape
    r, c = np.indices((H, W))
    lin = np.arange(H * W, dtype=np.int64).reshape(H, W)
    ring1 = (np.roll(f, 1, 0) + np.roll(f, -1, 0) + np.roll(f, 1, 1) + np.roll(f, -1, 1)) * (1.0 - f)
    ring2 = (np.roll(f, 2, 0) + np.roll(f, -2, 0) + np.roll(f, 2, 1) + np.roll(f, -2, 1)) * (1.0 - f)
    src1 = np.full((H, W), -1, np.int64)
    src2 = np.full((H, W), -1, np.int64)

    def inside(dr, dc, k):
        if dc == 1:
            return c + k < W - 1
        if dc == -1:
            return c - k > 0
        if dr == 1:
            return r + k < H - 1
        return r - k > 0

    def mask_at(dr, dc, k, ok):
        rr = np.clip(r + dr * k, 0, H - 1)
        cc = np.clip(c + dc * k, 0, W - 1)
        return ok & (f[rr, cc] > 0), lin[rr, cc]

    def visit(sel, near, unset):
        test = np.full((H, W), -1, np.int64)
        for dr, dc in _DIRS:
            cand = sel & inside(dr, dc, near) & (unset == -1)
            hit_near, lin_near = mask_at(dr, dc, near, cand)
            hit_far, lin_far = mask_at(dr, dc, near + 1, hit_near & inside(dr, dc, near + 1))
            src2[hit_far] = lin_far[hit_far]
            src1[hit_far] = lin_near[hit_far]
            single = hit_near & ~hit_far
            test[single] = lin_near[single]
        take = sel & (src1 == -1) & (test >= 0)
        src1[take] = test[take]

    visit(ring1 > 0, 1, src1)
    if two_pixel:
        visit((ring2 > 0) & (ring1 == 0), 2, src2)
    return src1.astype(np.int32), src2.astype(np.int32)


def resampling_map(input_shape, output_shape, rot_deg=0.0, row_shift=0.0, col_shift=0.0):
    """Affine map (output row Y, column X) -> input (y, x) of Interp2D's default grid:
    y = ay_y Y + ay_x X + y0, x = ax_y Y + ax_x X + x0, clamped to the input by the kernel.
    Returns the six coefficients (ay_y, ay_x, y0, ax_y, ax_x, x0)."""
    ry = (input_shape[0] - 1) / output_shape[0]
    rx = (input_shape[1] - 1) / output_shape[1]
    if rot_deg != 0:
        yc = input_shape[0] / 2 - 0.5
        xc = input_shape[1] / 2 - 0.5
        cs = math.cos(rot_deg * 3.1415 / 180.0)       # the reference's value of pi (interp2d.py:76)
        sn = math.sin(rot_deg * 3.1415 / 180.0)
        return (cs * ry, sn * rx, yc - sn * xc - cs * yc + row_shift,
                -sn * ry, cs * rx, xc - cs * xc + sn * yc + col_shift)
    return (ry, 0.0, float(row_shift), 0.0, rx, float(col_shift))
