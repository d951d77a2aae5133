"""A CPU model of the streaming inner-step pass (k_ppcg_inner_ring in
tealeaf_cuda_amd/csrc/engine_solver.hpp, whose per-cell arithmetic the other
pass forms share), checked against the oracle, and the hazard analysis of its
launch.

The model follows the kernel lane by lane: bands of 64 lanes (NS - 1 halo
lanes at the left, NS at the right), chunks of strips, the ring of NS + 1
strip slots, the fill-skip rule, lane shifts that give 0 at the end lanes,
loads that read 0 where the kernel's bounded buffers read 0, and every
expression in the kernel's order in the field type (numpy float32 / float64
round each operation as the GPU does without contraction).  Temporaries the
kernel leaves uninitialised (the Thomas sweep's rows past a short strip) are
NaN here, so a use of one would reach the compared outputs.

* the model's NS steps equal the oracle's NS single steps bit for bit for
  NS = 1..5, both precisions, the three preconditioners, ragged meshes, halo
  depths 1 and 2, and chunks of 1 strip up to the whole mesh;
* the first pass of an outer iteration that forms sd = M^-1 r / theta from r
  itself (block Jacobi) equals the oracle's z = M^-1 r, init_sd_new and NS
  steps bit for bit with NS halo lanes at the left and NS + 1 at the right,
  and not with NS at the right;
* utemp is read and written in place by the kernel.  A wave reads utemp at the
  strips of the neighbouring chunks it advances (written by other waves in the
  same launch), but those values never reach a store: the model gives the same
  outputs when every utemp value a wave reads outside its own chunk is NaN.
  sd and rtemp are read from one buffer and written to another, and every
  interior cell of the three outputs is written by exactly one wave.  So the
  launch has no cross-wave read-after-write hazard that can change a result,
  for any shape (the rectangle algebra of test_no_cross_wave_hazard covers
  4096^2 and 8192^2 at their launch shapes).
"""
from __future__ import annotations

import numpy as np
import pytest

from tealeaf_cuda_amd import Deck, Engine, StateSpec
from tealeaf_cuda_amd import runtime as rt

WAVE, STRIP = 64, 4
LAYOUT = [StateSpec(100.0, 0.0001), StateSpec(0.1, 25.0, "rectangle", 0.0, 3.0, 1.0, 6.0),
          StateSpec(3.0, 1.0, "rectangle", 5.0, 9.0, 2.0, 9.0)]


def _below(v):   # lane l gets lane l-1's value, lane 0 gets 0 (wave_shr:1, bound_ctrl)
    out = np.zeros_like(v)
    out[..., 1:] = v[..., :-1]
    return out


def _above(v):   # lane l gets lane l+1's value, the last lane gets 0 (wave_shl:1, bound_ctrl)
    out = np.zeros_like(v)
    out[..., :-1] = v[..., 1:]
    return out


def launch_shape(nx, ny, ns, per_simd, cu=256, chunk=None, from_r=False):
    """Bands, strips and chunk length as launch_inner_march chooses them."""
    ox = WAVE - 2 * ns + 1 - (2 if from_r else 0)
    nbands = (nx - 1 + ox) // ox
    nstrips = (ny - 1 + STRIP) // STRIP
    if chunk is None:
        nch = max(1, min(nstrips, cu * 4 * per_simd // nbands))
        chunk = (nstrips + nch - 1) // nch
    chunk = min(chunk, nstrips)
    return nbands, nstrips, chunk, (nstrips + chunk - 1) // chunk


def model_pass(SD, RT, UT, KX, KY, MI, h, nx, ny, ns, chunk, a, b, prec, poison_foreign_ut=False, from_r=False,
               theta=None, halo_r=None):
    """NS inner steps as the streaming pass computes them; returns the three
    output arrays (cells the pass does not write keep the input values) and
    the number of times each interior cell was written.  from_r (block
    Jacobi): the pass forms sd = M^-1 r / theta and utemp = sd inside the
    extent from RT (= r) as each strip enters, reading only sd's halo ring;
    halo_r overrides the right halo lanes (NS + 1 for that pass)."""
    T = SD.dtype.type
    x0, x1, y0, y1 = h, h + nx - 1, h, h + ny - 1
    hl, hr, rn = (ns, ns + 1, ns + 1) if from_r else (ns - 1, ns, ns + 1)
    if halo_r is not None:
        hr = halo_r
    ox = WAVE - hl - hr
    nbands, nstrips, chunk, nchunks = launch_shape(nx, ny, ns, 2, chunk=chunk, from_r=from_r)
    nbands = (nx - 1 + ox) // ox
    W = nbands * nchunks
    wid = np.arange(W)
    band, ck = wid % nbands, wid // nbands
    lane = np.arange(WAVE)[None, :]
    c = x0 + band[:, None] * ox - hl + lane                                  # (W, 64)
    col_in = (c >= x0) & (c <= x1)
    col_mem = (c >= x0 - 1) & (c <= x1 + 1)
    out_lane = (lane >= hl) & (lane <= WAVE - 1 - hr)
    s_first = ck * chunk
    s_last = np.minimum(s_first + chunk, nstrips) - 1
    i0, i1 = s_first - ns, s_last + ns
    rows, cols = SD.shape
    cc = np.clip(c, 0, cols - 1)
    SDo, RTo, UTo = SD.copy(), RT.copy(), UT.copy()
    writes = np.zeros((3, rows, cols), dtype=np.int32)

    def ld(A, r, cond):
        rr = np.clip(r, 0, rows - 1)
        return np.where(cond, A[rr, cc], T(0.0)).astype(T)

    def strip_len(r0):
        ln = np.where(r0 < y0, 0, y1 - r0 + 1)
        return np.clip(ln, 0, STRIP)

    def fetch(s):
        r0 = (y0 + s * STRIP)[:, None]
        ln = strip_len(r0)
        act = np.where(col_in, ln, 0)
        m = {"r0": r0, "len": ln}
        m["sd"] = [ld(SD, r0 + k, col_mem & (r0 + k >= y0 - 1) & (r0 + k <= y1 + 1) & ~(from_r & (k < act)))
                   for k in range(STRIP)]
        m["ky"] = [ld(KY, r0 + k, (k <= act) & (act > 0)) for k in range(STRIP + 1)]
        m["kx"] = [ld(KX, r0 + k, col_mem & (k < ln)) for k in range(STRIP)]
        m["rt"] = [ld(RT, r0 + k, k < act) for k in range(STRIP)]
        own = ((s >= s_first) & (s <= s_last))[:, None]
        ut = [ld(UT, r0 + k, (k < act) & out_lane & (not from_r)) for k in range(STRIP)]
        if poison_foreign_ut:   # values of other waves' chunks
            ut = [np.where(own | ~((k < act) & out_lane), u, T(np.nan)).astype(T) for k, u in enumerate(ut)]
        m["ut"] = ut
        m["mi"] = [ld(MI, r0 + k, k < act) if prec == "diag" else np.zeros((W, WAVE), T) for k in range(STRIP)]
        # lane 0's left neighbour column
        rr = np.clip(r0, 0, rows - 1)
        m["sde"] = [np.where((k < act) & (lane == 0) & (not from_r),
                             SD[np.clip(rr + k, 0, rows - 1), np.clip(c - 1, 0, cols - 1)], T(0.0)).astype(T)
                    for k in range(STRIP)]
        return m

    def apply_block(rv, diag, ky, ln, cp, bfp):   # strip_apply (rows past the strip stay NaN)
        zv = [np.full((W, WAVE), np.nan, T) for _ in range(STRIP)]
        dp = [np.full((W, WAVE), np.nan, T) for _ in range(STRIP)]
        dp[0] = rv[0] / diag[0]
        for k in range(1, STRIP):
            dp[k] = np.where(k < ln, (rv[k] - (-ky[k]) * dp[k - 1]) * bfp[k], T(np.nan)).astype(T)
        for k in range(STRIP - 1, -1, -1):
            if k == STRIP - 1:
                zv[k] = np.where(ln == STRIP, dp[k], zv[k]).astype(T)
            else:
                z_last = np.where(ln == k + 1, dp[k], zv[k])
                zv[k] = np.where(k < ln - 1, dp[k] - cp[k] * zv[k + 1], z_last).astype(T)
        return zv

    def factor(m):
        ln = m["len"]
        cp = [np.zeros((W, WAVE), T) for _ in range(STRIP)]
        bfp = [np.zeros((W, WAVE), T) for _ in range(STRIP)]
        if prec == "block":
            diag = [T(1.0) + (m["ky"][k + 1] + m["ky"][k]) + (_above(m["kx"][k]) + m["kx"][k]) for k in range(STRIP)]
            with np.errstate(all="ignore"):
                cp[0] = np.where(ln > 0, (-m["ky"][1]) / diag[0], cp[0]).astype(T)
                for k in range(1, STRIP):
                    bk = T(1.0) / (diag[k] - (-m["ky"][k]) * cp[k - 1])
                    ck_ = (-m["ky"][k + 1]) * bk
                    bfp[k] = np.where(k < ln, bk, T(0.0)).astype(T)
                    cp[k] = np.where(k < ln, ck_, T(0.0)).astype(T)
                if from_r:   # sd = utemp = M^-1 r / theta inside the extent (the last lane lacks kxr: stale)
                    zv = apply_block(m["rt"], diag, m["ky"], ln, cp, bfp)
                    theta_r = T(1.0) / T(theta)
                    for k in range(STRIP):
                        inside = col_in & (k < ln)
                        m["sd"][k] = np.where(inside, zv[k] * theta_r, m["sd"][k]).astype(T)
                        m["ut"][k] = np.where(inside, m["sd"][k], m["ut"][k]).astype(T)
        m["cp"], m["bfp"] = cp, bfp

    def advance(m, lo, hi, sde, st, active):
        al, be = T(a[st]), T(b[st])
        ln = m["len"]
        sd, kx, ky = m["sd"], m["kx"], m["ky"]
        rv, diag = [], []
        with np.errstate(all="ignore"):
            for k in range(STRIP):
                fl, fr, kxr = _below(sd[k]), _above(sd[k]), _above(kx[k])
                dg = T(1.0) + (ky[k + 1] + ky[k]) + (kxr + kx[k])
                below = lo if k == 0 else sd[k - 1]
                above = hi if k == STRIP - 1 else sd[k + 1]
                left = fl if st > 0 else np.where(lane > 0, fl, sde[k])
                res = dg * sd[k] - (ky[k + 1] * above + ky[k] * below) - (kxr * fr + kx[k] * left)
                rv.append((m["rt"][k] - res).astype(T))
                diag.append(dg)
            zv = [np.full((W, WAVE), np.nan, T) for _ in range(STRIP)]
            if prec == "block":   # strip_apply over the strip's rows (rows past it stay NaN)
                zv = apply_block(rv, diag, ky, ln, m["cp"], m["bfp"])
            for k in range(STRIP):
                if prec == "block":
                    sdn = al * sd[k] + be * zv[k]
                elif prec == "diag":
                    sdn = al * sd[k] + be * rv[k] * m["mi"][k]
                else:
                    sdn = al * sd[k] + be * rv[k]
                upd = active[:, None] & (k < ln)
                m["sd"][k] = np.where(upd & col_in, sdn, sd[k]).astype(T)
                m["rt"][k] = np.where(upd, rv[k], m["rt"][k]).astype(T)
                m["ut"][k] = np.where(upd, m["ut"][k] + sdn, m["ut"][k]).astype(T)

    zero = lambda: [np.zeros((W, WAVE), T) for _ in range(STRIP)]   # noqa: E731
    R = [{"sd": zero(), "rt": zero(), "ut": zero(), "kx": zero(), "ky": zero() + [np.zeros((W, WAVE), T)],
          "cp": zero(), "bfp": zero(), "mi": zero(), "r0": np.full((W, 1), y0), "len": np.zeros((W, 1), int)}
         for _ in range(rn)]
    P = [np.zeros((W, WAVE), T) for _ in range(ns)]
    sde0 = [np.zeros((W, 1), T) for _ in range(STRIP)]
    nxt = fetch(i0)
    tmax = int((i1 - i0).max())
    for t in range(tmax + 1):
        i = i0 + t
        live = i <= i1
        ph = t % rn
        cur = nxt
        factor(cur)
        R[ph] = {k: v for k, v in cur.items() if k != "sde"}
        sde_cur = [cur["sde"][k][:, 0:1] for k in range(STRIP)]   # lane 0's values, broadcast
        nxt = fetch(i + 1)
        hi = R[ph]["sd"][0]
        for m in range(ns):
            j = (ph - 1 - m) % rn
            hslot = R[j]
            last = hslot["sd"][STRIP - 1]
            s = i - 1 - m
            dist = np.where(s < s_first, s_first - s, np.where(s > s_last, s - s_last, 0))
            active = live & (dist <= ns - 1 - m) & (hslot["len"][:, 0] > 0)
            advance(hslot, P[m], hi, sde0, m, active)
            P[m] = last
            hi = hslot["sd"][0]
        f = R[(ph + 1) % rn]
        sf = i - ns
        out = (out_lane & col_in & (live & (sf >= s_first) & (sf <= s_last))[:, None])
        for k in range(STRIP):
            wmask = out & (k < f["len"])
            ww, ll = np.nonzero(wmask)
            rr = (f["r0"][:, 0] + k)[ww]
            cl = c[ww, ll]
            SDo[rr, cl] = f["sd"][k][ww, ll]
            RTo[rr, cl] = f["rt"][k][ww, ll]
            UTo[rr, cl] = f["ut"][k][ww, ll]
            for q in range(3):
                np.add.at(writes[q], (rr, cl), 1)
        sde0 = sde_cur
    return SDo, RTo, UTo, writes


def _oracle_case(oracle_so, tmp_path, nx, ny, h, prec, fp32, seed):
    eng = Engine(oracle_so)
    pname = {"block": "jac_block", "diag": "jac_diag", "none": "none"}[prec]
    deck = Deck(x_cells=nx, y_cells=ny, solver="ppcg", preconditioner=pname, halo_depth=h, states=LAYOUT)
    eng.initialise(deck.write(tmp_path / "m.in"), nx, ny)
    eng.initialise_chunk(0.0, 0.0, 10.0 / nx, 10.0 / ny)
    st = LAYOUT
    eng.generate_chunk([s.density for s in st], [s.energy for s in st], [s.xmin for s in st], [s.xmax for s in st],
                       [s.ymin for s in st], [s.ymax for s in st], [s.radius for s in st],
                       [{"rectangle": 1}.get(s.geometry or "", 0) for s in st])
    eng.set_field_kernel()
    eng.common_init(rt.CONDUCTIVITY, 0.04)
    if fp32:
        eng.set_precision(32)
        eng.mp_setup()
    eng.cg_init()
    get = eng.get_field32 if fp32 else eng.get_field
    put = eng.set_field32 if fp32 else eng.set_field
    dt = np.float32 if fp32 else np.float64
    rng = np.random.default_rng(seed)
    fields = {f: rng.standard_normal(eng.shape()).astype(dt) for f in (rt.FIELD_SD, rt.XFIELD_RTEMP, rt.XFIELD_UTEMP)}
    for f, v in fields.items():
        put(f, v)
    coef = dict(KX=get(rt.XFIELD_KX), KY=get(rt.XFIELD_KY), MI=get(rt.XFIELD_MI))
    return eng, fields, coef, get


CASES = [  # nx, ny, halo, ns, chunk, preconditioner
    (96, 80, 1, 5, 3, "block"), (96, 80, 1, 5, None, "block"), (131, 77, 1, 4, 2, "block"),
    (131, 77, 2, 3, 7, "diag"), (70, 45, 2, 2, 1, "none"), (97, 66, 1, 1, 4, "block"),
    (150, 130, 1, 5, 40, "diag"), (64, 12, 1, 3, 1, "none"), (57, 9, 2, 4, 2, "block"),
]


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("nx,ny,h,ns,chunk,prec", CASES)
def test_model_equals_oracle(nx, ny, h, ns, chunk, prec, fp32, oracle_so, oracle_env, tmp_path):
    oracle_env("dot2")
    a = [0.3, 0.55, 0.7, 0.2, 0.6, 0.45, 0.5]
    b = [1.1, 0.9, 0.8, 1.2, 0.7, 1.0, 0.95]
    eng, fields, coef, get = _oracle_case(oracle_so, tmp_path, nx, ny, h, prec, fp32, seed=nx + ns)
    try:
        eng.ppcg_init_constants(a, b)
        eng.ppcg_inner_n(1, ns)
        expect = [get(f) for f in (rt.FIELD_SD, rt.XFIELD_RTEMP, rt.XFIELD_UTEMP)]
    finally:
        eng.finalize()
    SD, RT, UT = (fields[f] for f in (rt.FIELD_SD, rt.XFIELD_RTEMP, rt.XFIELD_UTEMP))
    got = model_pass(SD, RT, UT, coef["KX"], coef["KY"], coef["MI"], h, nx, ny, ns, chunk, a, b, prec)
    inner = (slice(h, h + ny), slice(h, h + nx))
    for g_, e_, name in zip(got[:3], expect, ("sd", "rtemp", "utemp")):
        assert np.array_equal(g_[inner], e_[inner]), f"{name}: {np.sum(g_[inner] != e_[inner])} cells differ"
    writes = got[3]
    assert np.all(writes[:, inner[0], inner[1]] == 1)          # every interior cell once, per output
    assert writes.sum() == 3 * nx * ny                          # nothing outside the interior
    # the utemp values a wave reads outside its own chunk never reach a store
    poisoned = model_pass(SD, RT, UT, coef["KX"], coef["KY"], coef["MI"], h, nx, ny, ns, chunk, a, b, prec,
                          poison_foreign_ut=True)
    for g_, p_ in zip(got[:3], poisoned[:3]):
        assert np.array_equal(g_[inner], p_[inner])


@pytest.mark.parametrize("n,ns,per_simd", [(4096, 5, 2), (4096, 4, 2), (4096, 3, 1), (4096, 4, 1),
                                           (8192, 5, 2), (8192, 3, 1), (1024, 5, 2), (1000, 4, 2)])
def test_no_cross_wave_hazard(n, ns, per_simd):
    """The launch shapes at full size, by rectangle algebra: the rows each
    wave writes (its chunk) partition the interior rows per band, the bands'
    written columns partition the interior columns, and a wave's utemp reads
    outside its own chunk fall on rows other waves write, which the model
    shows never reach a store; sd and rtemp are written to other buffers."""
    h, nx, ny = 1, n, n
    x0, x1, y0, y1 = h, h + nx - 1, h, h + ny - 1
    ox = WAVE - 2 * ns + 1
    nbands, nstrips, chunk, nchunks = launch_shape(nx, ny, ns, per_simd)
    cols_written = []
    for band in range(nbands):
        c0 = x0 + band * ox - (ns - 1)
        lo, hi = max(c0 + ns - 1, x0), min(c0 + WAVE - 1 - ns, x1)   # out lanes inside the extent
        cols_written.append((lo, hi))
    assert cols_written[0][0] == x0 and cols_written[-1][1] == x1
    assert all(cols_written[k][1] + 1 == cols_written[k + 1][0] for k in range(nbands - 1))
    rows_written = []
    for ck in range(nchunks):
        sf, sl = ck * chunk, min(ck * chunk + chunk, nstrips) - 1
        rows_written.append((y0 + sf * STRIP, min(y0 + sl * STRIP + STRIP - 1, y1)))
        # utemp rows this wave reads: its strips and ns on each side, inside the extent
        read = (max(y0, y0 + (sf - ns) * STRIP), min(y1, y0 + (sl + ns) * STRIP + STRIP - 1))
        own = rows_written[-1]
        foreign = [(read[0], own[0] - 1), (own[1] + 1, read[1])]
        for lo, hi in foreign:
            if lo <= hi:   # rows of the neighbouring chunks only
                assert lo >= y0 and hi <= y1 and (hi < own[0] or lo > own[1])
                assert own[0] - lo <= ns * STRIP and hi - own[1] <= ns * STRIP
    assert rows_written[0][0] == y0 and rows_written[-1][1] == y1
    assert all(rows_written[k][1] + 1 == rows_written[k + 1][0] for k in range(nchunks - 1))


FROM_R_CASES = [(96, 80, 1, 5, 3), (96, 80, 1, 5, None), (131, 77, 1, 4, 2), (57, 9, 2, 4, 2), (150, 130, 1, 5, 40)]


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("nx,ny,h,ns,chunk", FROM_R_CASES)
def test_from_r_model_equals_oracle(nx, ny, h, ns, chunk, fp32, oracle_so, oracle_env, tmp_path):
    """The first pass of an outer iteration with sd formed from r
    (k_ppcg_inner_ringl FROM_R, block Jacobi): the model's pass over r equals
    the oracle's z = M^-1 r (ppcg_init step 2) and init_sd_new(theta)
    followed by NS steps, bit for bit, with NS halo lanes at the left and
    NS + 1 at the right.  With NS at the right it does not: the last lane
    forms its own sd without its right neighbour's Kx, and that stale value
    reaches a written column in NS steps."""
    oracle_env("dot2")
    a = [0.3, 0.55, 0.7, 0.2, 0.6, 0.45, 0.5]
    b = [1.1, 0.9, 0.8, 1.2, 0.7, 1.0, 0.95]
    theta = 1.7
    eng, fields, coef, get = _oracle_case(oracle_so, tmp_path, nx, ny, h, "block", fp32, seed=3 * nx + ns)
    put = eng.set_field32 if fp32 else eng.set_field
    R = np.random.default_rng(nx * ny).standard_normal(eng.shape()).astype(np.float32 if fp32 else np.float64)
    try:
        put(rt.FIELD_R, R)
        eng.ppcg_init_constants(a, b)
        eng.ppcg_init(2)
        eng.ppcg_init_sd_new(theta)
        eng.ppcg_inner_n(1, ns)
        expect = [get(f) for f in (rt.FIELD_SD, rt.XFIELD_RTEMP, rt.XFIELD_UTEMP)]
    finally:
        eng.finalize()
    SD, UT = fields[rt.FIELD_SD], fields[rt.XFIELD_UTEMP]
    args = (SD, R, UT, coef["KX"], coef["KY"], coef["MI"], h, nx, ny, ns, chunk, a, b, "block")
    got = model_pass(*args, from_r=True, theta=theta)
    inner = (slice(h, h + ny), slice(h, h + nx))
    for g_, e_, name in zip(got[:3], expect, ("sd", "rtemp", "utemp")):
        assert np.array_equal(g_[inner], e_[inner]), f"{name}: {np.sum(g_[inner] != e_[inner])} cells differ"
    assert np.all(got[3][:, inner[0], inner[1]] == 1) and got[3].sum() == 3 * nx * ny
    if nx >= WAVE - ns:   # the first band's last lane lies inside the extent
        short = model_pass(*args, from_r=True, theta=theta, halo_r=ns)
        assert not np.array_equal(short[0][inner], expect[0][inner])
