"""Closed-form models used by the fine-arithmetic pins of the oracle.

Each function computes, in float64 and from the geometry, what a stage of
the reference must produce, without reusing the oracle's fp32 code paths:

* shader_texel: the texel colorize_buddha_soft.frag:12-77 samples for each
  screen pixel (GL_NEAREST = floor of the texture coordinate times the
  texture size, GL_CLAMP_TO_EDGE), written with the shader's own step()
  weights;
* column_span: for a scene holding one voxel column, the texels, colour
  indices and depths render_line (Cuda_Render.h:456-733) draws on one ray
  line, from the exact entry and exit distances of the line's horizontal ray
  through the column's cell (slab method), the projection of the column's
  edges and the perspective-correct interpolation along its near edge
  (ray/segment intersection).

Both return, beside the expected values, a mask of the results that sit
within a small distance of a rounding boundary, where the fp32 chain of the
reference may land on either side.
"""
import math

import numpy as np


def step(edge, x):
    """GLSL step(edge, x): 0 where x < edge, else 1."""
    return np.where(x < edge, 0.0, 1.0)


def shader_texel(rm, rel=1e-6):
    """(row, column) texel each pixel of the W x H screen samples, row 0 at
    the top of the image; -1 where the coordinate is not finite.  `near`
    marks pixels whose texture coordinate lies within `rel` of the texture
    size (scaled by the size of the intermediate terms) of a texel edge."""
    W, H, rs = rm.W, rm.H, rm.rs
    RESX, RESY = float(W), float(H)
    rows, cols = np.mgrid[0:H, 0:W]
    fcx = cols + 0.5
    fcy = (H - 1 - rows) + 0.5
    vx, vy = float(rm.vanish[0]), float(rm.vanish[1])
    ofs = [float(v) for v in rm.ofs_add]
    rxg = float(rm.rot_x_greater_zero)
    ratio = float(rm.ray_ratio)
    scx, scy = fcx / RESX, fcy / RESY
    border = (RESX - RESY) / (RESX * 2)
    scx1, scy1 = scx - vx, scy - vy
    upper = step(scy1, 0.0)
    left = step(scx1, 0.0)
    ostep = step(np.abs(scy1) - np.abs(scx1) * RESX / RESY, 0.0)
    seg_up = (1 - upper) * (1 - ostep)
    seg_dn = upper * (1 - ostep)
    seg_rt = (1 - left) * ostep
    seg_lt = left * ostep
    o2 = (ostep * fcx + (1 - ostep) * fcy) / RESX
    with np.errstate(divide="ignore", invalid="ignore"):
        ang2 = scx1 * np.abs(1 - upper - vy) / scy1 + upper * (1 - vx) + (1 - upper) * vx
        ang3 = scy1 * np.abs(1 - left - vx) / scx1 + left * (1 - vy) + (1 - left) * vy
    ang3 = ang3 * RESY / RESX + border
    # the step() weights are exactly 0 or 1: select, so an infinite term of
    # the other branch does not turn the sum into NaN
    x_pre = np.where(ostep > 0, ang3, ang2)
    ty = np.select([seg_dn > 0, seg_up > 0, seg_lt > 0, seg_rt > 0],
                   [ofs[1] + x_pre, ofs[0] + 1.0 - x_pre, ofs[3] + x_pre, ofs[2] + 1.0 - x_pre])
    ty = ty * ratio * 0.25
    seg_up_x = rxg * seg_up + (1 - rxg) * seg_dn
    seg_dn_x = rxg * seg_dn + (1 - rxg) * seg_up
    seg_rt_x = rxg * seg_rt + (1 - rxg) * seg_lt
    seg_lt_x = rxg * seg_lt + (1 - rxg) * seg_rt
    tx = seg_up_x * (o2 + border) + seg_dn_x * (1.0 - (o2 + border)) + seg_rt_x * o2 \
        + seg_lt_x * (1.0 - o2)
    finite = np.isfinite(tx) & np.isfinite(ty)
    u = np.where(finite, tx * rs, 0.0)
    v = np.where(finite, ty * 4 * rs, 0.0)
    # an fp32 evaluation of the same terms errs by a few ulp of its largest
    # intermediate (|x_pre| grows near the vanishing point)
    scale = np.where(finite, 1.0 + np.abs(x_pre) + np.abs(ty), 1.0)
    near_u = np.abs(u - np.round(u)) < rel * rs * scale
    near_v = np.abs(v - np.round(v)) < rel * 4 * rs * scale
    tc = np.clip(np.floor(u), 0, rs - 1).astype(np.int64)
    tr = np.clip(np.floor(v), 0, 4 * rs - 1).astype(np.int64)
    tc[~finite] = -1
    tr[~finite] = -1
    return tr, tc, (near_u | near_v) & finite, finite


def trunc(v):
    return math.trunc(v)


def column_span(rm, plane, clip, cell, top, height, tol_px=0.02, tol_val=0.01):
    """What one ray line draws of a lone column.

    plane: (ray_x, ray_z, vertical) of the line; clip: its cleared range
    [lo, hi]; cell: (cx, cz) column cell; the column holds `height` voxels
    from row `top` down, colour indices 0..height-1 from the top.

    Returns None when the line's ray misses the cell or grazes its edges
    (ambiguous), else dict(ys, u, depth, ambiguous) for the texels drawn:
    ys the texel indices, u the colour index and depth the depth word of
    each, ambiguous a mask of texels whose u or depth sits near a rounding
    boundary; `uncertain_span` is set when an end of the span sits near a
    pixel boundary.
    """
    ray_x, ray_z, vertical = float(plane[0]), float(plane[1]), bool(plane[2])
    sin_x, cos_x = float(rm.sin_x), float(rm.cos_x)
    sin_y, cos_y = float(rm.sin_y), float(rm.cos_y)
    vpx, vpy, vpz = (float(v) for v in rm.position)
    rs = rm.rs
    res_x2 = float(rs // 2)
    res_y2 = float(rs // 2)
    # world direction of the horizontal ray (Cuda_Render.h:271-272)
    dx = ray_x * cos_y + ray_z * sin_y
    dz = ray_x * sin_y - ray_z * cos_y
    cx, cz = cell
    tmin, tmax = -math.inf, math.inf
    for o, d, c in ((vpx, dx, cx), (vpz, dz, cz)):
        if d == 0.0:
            if not (c < o < c + 1):
                return None
            continue
        t0, t1 = (c - o) / d, (c + 1 - o) / d
        if t0 > t1:
            t0, t1 = t1, t0
        tmin, tmax = max(tmin, t0), min(tmax, t1)
    if tmin >= tmax or tmin <= 0:
        return None
    if tmax - tmin < 1e-3:  # through a corner: the DDA may step either side
        return "ambiguous"
    # orientation of the line's texel axis (Cuda_Render.h:203-212)
    reverse = (vertical and ray_z <= 0) or (not vertical and ray_x <= 0 and sin_x > 0) or \
        (not vertical and ray_x > 0 and sin_x < 0)
    rxmr = -res_x2 if reverse else res_x2
    if vertical:
        rxmr = -rxmr

    def ZY(t, h):
        """Camera-plane depth and offset of the point at distance t along the
        line's horizontal ray, row h of the volume."""
        z = cos_x * (ray_z * t) + sin_x * (vpy + h)
        y = (cos_x * (vpy + h) - sin_x * (ray_z * t)) if vertical else ray_x * t
        return z, y * rxmr

    lo, hi = int(clip[0]), int(clip[1])
    a, b = float(top), float(top + height)
    # clip-top cull at row 0 of the near edge (Cuda_Render.h:456-467)
    z0, y0 = ZY(tmin, 0.0)
    if z0 * res_y2 + y0 <= z0 * lo:
        if abs((z0 * res_y2 + y0) - z0 * lo) < 1e-4 * max(1.0, abs(z0) * rs):
            return "ambiguous"
        return {"ys": np.zeros(0, int), "u": np.zeros(0), "depth": np.zeros(0), "ambiguous":
                np.zeros(0, bool), "uncertain_span": False}
    # span ends: the top at the far edge when it lies below the eye, the
    # bottom at the far edge when it lies above it
    zt, yt = ZY(tmax if vpy + a >= 0 else tmin, a)
    zb, yb = ZY(tmax if vpy + b < 0 else tmin, b)
    empty = {"ys": np.zeros(0, int), "u": np.zeros(0), "depth": np.zeros(0),
             "ambiguous": np.zeros(0, bool), "uncertain_span": False}
    if zt <= 0 or zb <= 0:
        return empty
    s2 = res_y2 + yt / zt
    s1 = res_y2 + yb / zb - 1.0
    uncertain = abs(s2 - round(s2)) < tol_px or abs(s1 - round(s1)) < tol_px
    scr_y2, scr_y1 = trunc(s2), trunc(s1)
    if scr_y2 <= lo or scr_y1 >= hi:
        empty["uncertain_span"] = uncertain
        return empty
    scr_y2 = min(scr_y2, hi)
    scr_y1 = max(scr_y1, lo)
    ys = np.arange(scr_y1, scr_y2)
    # perspective-correct interpolation along the near edge: the point of
    # the segment bottom(u = height) -> top(u = 0) whose projection is y + 1
    Bz, By = ZY(tmin, b)
    Tz, Ty = ZY(tmin, a)
    q = ys + 1.0 - res_y2
    lam = (q * Bz - By) / ((Ty - By) - q * (Tz - Bz))
    u = height * (1.0 - lam)
    z = Bz + lam * (Tz - Bz)
    ui = np.clip(np.trunc(u), 0, height - 1).astype(np.int64)
    di = (np.trunc(z).astype(np.int64)) & ~1
    amb = (np.abs(u - np.round(u)) < tol_val * max(1.0, height / 64)) & (u > 0) & (u < height)
    amb |= np.abs(z - np.round(z)) < tol_val * np.maximum(1.0, z / 64)
    return {"ys": ys, "u": ui, "depth": di, "ambiguous": amb, "uncertain_span": uncertain}


def clip_ranges(rm, planes):
    """Cleared texel range of every ray line (Cuda_Render.h:203-250), from
    the line's clipped 2D end points: each end point scaled to pixels (one
    fp32 product, as the source) and truncated, moved outward by 2 pixels
    (inward by 1 where the line's texel axis runs against the screen axis),
    clamped to the screen, mapped to the texel axis and sorted.  Lines whose
    end points coincide on the screen clear nothing: (0, -1)."""
    rs = rm.rs
    n = len(planes)
    out = np.zeros((n, 2), np.int64)
    sin_x = float(rm.sin_x)
    f_rs = np.float32(rs)
    for L in range(n):
        ray_x, ray_z, vertical = float(planes[L, 0]), float(planes[L, 1]), planes[L, 2] > 0
        sx, sy, ex, ey = (float(f_rs * np.float32(v)) for v in planes[L, 3:7])
        reverse = (vertical and ray_z <= 0) or (not vertical and ray_x <= 0 and sin_x > 0) or \
            (not vertical and ray_x > 0 and sin_x < 0)
        pad = 1 if reverse else -2
        p1 = [min(max(math.trunc(v) + pad, 0), rs - 1) for v in (sx, sy)]
        p2 = [min(max(math.trunc(v) - pad, 0), rs - 1) for v in (ex, ey)]
        if p1[1] == p2[1]:
            out[L] = (0, -1)
            continue
        k = 1 if vertical else 0
        lo, hi = rs - 1 - p1[k], rs - 1 - p2[k]
        if reverse:
            lo, hi = rs - 1 - lo, rs - 1 - hi
        lo, hi = min(lo, hi), max(lo, hi)
        out[L] = (lo, hi) if lo < hi else (0, -1)
    return out
