"""The eigenvalue sweeps leave out the evanescent top of the grid (csrc/sturm.h evanescent_rows).

For a block of 64 consecutive eigenvalues whose shifts all lie above s_min, rows above a cut
n_eff are dropped from every sweep and n - n_eff is added to each count.  Two claims carry
that, checked here on the full 20000-row matrices of the synthetic profile family with
scipy's LAPACK (dstebz) as the independent side:
  * counts: the full matrix's Sturm count at any shift s >= s_min equals the truncated
    matrix's count plus n - n_eff (integer equality, shifts right next to eigenvalues too);
  * eigenvalues: the truncated matrix's eigenvalues in the block's range equal the full
    matrix's to rounding (a few ulp of ||dz^2 T||), far inside the 1e-9 parity bound.
The cut rule is restated in numpy from the kernel's description (same thresholds).
"""
import math

import numpy as np
import pytest
import scipy.linalg

from oracle import ess
from oracle.atmos import AtmosphereNodes, read_profile

DECAY = 80.0


def cut_rows(a1, s_min, decay=DECAY, min_keep=0):
    """n_eff for the T' diagonal a1 (off-diagonal 1) at shifts >= s_min."""
    n = a1.size
    evan = (a1 - s_min) <= -2.0
    osc = np.nonzero(~evan)[0]
    t0 = int(osc[-1]) + 1 if osc.size else 0
    y = 0.5 * (-2.0 - (a1[t0:] - s_min))
    d = 2.0 * np.log1p(y + np.sqrt(y * (y + 2.0)))
    hit = np.nonzero(np.cumsum(d) >= decay)[0]
    n_eff = t0 + int(hit[0]) + 1 if hit.size else n
    return min(n, max(n_eff, min_keep))


def count_T1(a1, s):
    """Eigenvalues of tridiag(1, a1, 1) below each shift in s (LDL^T inertia from the top, as
    the sweeps run; a zero pivot is replaced by a tiny negative one, as LAPACK does)."""
    s = np.asarray(s, dtype=np.float64)
    d = a1[-1] - s
    cnt = (d < 0).astype(np.int64)
    for i in range(a1.size - 2, -1, -1):
        d = (a1[i] - s) - 1.0 / np.where(d == 0.0, -1e-300, d)
        cnt += d < 0
    return cnt


def problem(path, freq, az, Nz=20000):
    nodes = AtmosphereNodes(read_profile(path))
    grid = ess.EssGrid(nodes, Nz=Nz)
    diag, _, k_min, k_max, _ = ess.modal_trace(grid, freq, az, 0.0, 0.0, True)
    a, e = ess.matrix(diag, grid.dz)
    h2 = grid.dz ** 2
    return a * h2, h2, k_min, k_max


@pytest.fixture(scope="module")
def profile(tmp_path_factory):
    from ncpamodes import synth
    d = tmp_path_factory.mktemp("evan")
    return synth.write_profile(str(d / "P3.dat"), seed=3)


@pytest.mark.parametrize("freq,az", [(0.5, 90.0), (1.0, 270.0), (2.0, 0.0)])
def test_cut_keeps_counts_and_eigenvalues(profile, freq, az):
    a1, h2, k_min, k_max = problem(profile, freq, az)
    n = a1.size
    ones = np.ones(n - 1)
    w = scipy.linalg.eigh_tridiagonal(a1, ones, eigvals_only=True, select="v",
                                      select_range=(h2 * k_min ** 2, h2 * k_max ** 2), lapack_driver="stebz")
    assert w.size > 64
    rng = np.random.default_rng(7)
    dropped = 0
    blocks = list(range(0, w.size, 64))
    for b in blocks[1::2] + blocks[:1]:   # every other block and the lowest one
        blk = w[b:b + 64]
        s_min = blk[0] - 1e-9 * (blk[-1] - blk[0] + 1e-12)     # a lower bound of the block
        n_eff = cut_rows(a1, s_min)
        dropped += n - n_eff
        if n_eff == n:
            continue
        # left-out rows: every one evanescent at s_min
        assert np.all(a1[n_eff:] - s_min <= -2.0)
        # counts at shifts right next to the block's eigenvalues and at random shifts
        gap = np.min(np.diff(blk)) if blk.size > 1 else 1e-6
        shifts = np.concatenate([blk[:8] - 1e-3 * gap, blk[:8] + 1e-3 * gap,
                                 rng.uniform(s_min, blk[-1] + gap, 4)])
        full, trunc = count_T1(a1, shifts), count_T1(a1[:n_eff], shifts)
        assert np.array_equal(full, trunc + (n - n_eff)), (b, full - trunc)
        # eigenvalues of the truncated matrix in the block's range
        wt = scipy.linalg.eigh_tridiagonal(a1[:n_eff], ones[:n_eff - 1], eigvals_only=True, select="v",
                                           select_range=(blk[0] - 0.5 * gap, blk[-1] + 0.5 * gap),
                                           lapack_driver="stebz")
        assert wt.size == blk.size
        tol = 8 * np.finfo(float).eps * 4.0
        assert np.max(np.abs(wt - blk)) <= tol, np.max(np.abs(wt - blk))
    if freq >= 1.0:
        assert dropped > 0   # the cut does leave rows out at these frequencies


def test_cut_is_conservative_without_evanescent_top(profile):
    # a shift below every row's a' + 2 leaves nothing out; so does a top that is not evanescent
    a1, h2, k_min, k_max = problem(profile, 0.5, 90.0)
    assert cut_rows(a1, float(np.min(a1)) - 1.0) == a1.size
    b = a1.copy()
    b[-1] = 10.0
    assert cut_rows(b, h2 * k_max ** 2) == a1.size
    # records below the cut are kept
    n_eff = cut_rows(a1, h2 * k_max ** 2)
    assert n_eff < a1.size
    assert cut_rows(a1, h2 * k_max ** 2, min_keep=a1.size - 3) == a1.size - 3


def _cut_rows_margin(am, decay, margin):
    """The rule on row values am at the bound (the wide-angle pencil's a_i(k_min) = A + k (B +
    k C)), rows with am <= -2 - margin evanescent."""
    n = am.size
    osc = np.nonzero(~(am <= -2.0 - margin))[0]
    t0 = int(osc[-1]) + 1 if osc.size else 0
    y = 0.5 * (-2.0 - am[t0:])
    d = 2.0 * np.log1p(y + np.sqrt(y * (y + 2.0)))
    hit = np.nonzero(np.cumsum(d) >= decay)[0]
    return t0 + int(hit[0]) + 1 if hit.size else n


def test_wide_angle_cut(profile):
    """The wide-angle pencil dz^2 T(k) = tridiag(1, A + k (B + k C), 1) with k as the shift
    (csrc/wide.hip WidePencil::narrow): a_i(k) + 2 is concave in k with a positive larger root
    (subsonic winds), so rows evanescent at k_min stay evanescent above it; full counts equal
    truncated counts plus the dropped rows, and QEP eigenvalues of the truncated pencil equal
    the full ones (oracle/wmod.py's branch crossings) to working precision."""
    import scipy.optimize
    from oracle import wmod
    nodes = AtmosphereNodes(read_profile(profile))
    grid = wmod.WGrid(nodes, Nz=20000)
    kd, md, cd, _, _, k_min, k_max, _ = wmod.modal_trace(grid, 1.0, 270.0, 0.0, 0.0, True)
    A, B, C = wmod.pencil(kd, md, cd, grid.dz_m)
    n = A.size
    assert np.all(md < 0)
    Pf = (A, B, C)
    n_lo, n_hi = wmod._n_negative(Pf, k_min), wmod._n_negative(Pf, k_max)
    assert n_hi - n_lo > 200

    def count(Pm, ks):
        a = Pm[0][:, None] + ks[None, :] * (Pm[1][:, None] + ks[None, :] * Pm[2][:, None])
        d = a[-1] - 0.0
        cnt = (d < 0).astype(np.int64)
        for i in range(a.shape[0] - 2, -1, -1):
            d = a[i] - 1.0 / np.where(d == 0.0, -1e-300, d)
            cnt += d < 0
        return cnt

    def branch(Pm, i):
        return scipy.optimize.brentq(lambda k: wmod._mu(Pm, k, i), k_min, k_max, xtol=1e-300,
                                     rtol=4 * np.finfo(float).eps, maxiter=200)

    rng = np.random.default_rng(3)
    tested = 0
    for b in (64, 192, n_hi - n_lo - 64):           # blocks of 64 branches
        k0 = branch(Pf, n_lo + b)
        k_lo = k0 * (1 - 1e-12)
        n_eff = _cut_rows_margin(A + k_lo * (B + k_lo * C), DECAY, 1e-12)
        if n_eff == n:
            continue
        tested += 1
        Pt = (A[:n_eff], B[:n_eff], C[:n_eff])
        k1 = branch(Pf, n_lo + b + 1)
        ks = np.concatenate([[k0 * (1 + 1e-10), k1 * (1 - 1e-10)], rng.uniform(k_lo, k_max, 4)])
        assert np.array_equal(count(Pf, ks), count(Pt, ks) + (n - n_eff))
        for i in (n_lo + b, n_lo + b + 1):
            kf, kt = branch(Pf, i), branch(Pt, i - (n - n_eff))
            assert abs(kt - kf) <= 1e-14 * kf, (i, kf, kt)
    assert tested >= 1
