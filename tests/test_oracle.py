"""Pins the CPU oracle against what the mathematics fixes.

The reference has no tests and no golden vectors (its only fixture is the
overthrust model).  Each oracle function is pinned here by closed forms,
invariants or brute force, chosen so that a dropped term, a wrong sign or
index, or a transposed operand fails:

  stencil      polynomial exactness, the Fornberg weights, the Fourier symbol
  forward      linearity, discrete reciprocity, zero rim, sponge energy decay,
               receiver / injection timing against a brute-force statement
  reverse      exact time reversal of the leapfrog (illumination identity),
               image = sum of S*R against forward fields
  wavelet      DFT pair identity, a library FFT (numpy's pocketfft) at the
               wavelet lengths, filter modulus sqrt(2 pi f) and phase pi/4
  sponge       closed form of the Cerjan table and the corner rule
  pad / index  brute force against numpy edge padding, half-cell rounding
  post / step  closed forms of -d2/dz2 and of the parabola vertex
  random bdry  an independent statement of the generator, bit for bit
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from cases import layered, rel_l2

FORNBERG = [-205.0 / 72, 8.0 / 5, -1.0 / 5, 8.0 / 315, -1.0 / 560]
# the reference's decimal weights FDM5..FDM1 (src/modeling/modeling.cuh:9-13)
REF_WEIGHTS = [-2.847222222222, 1.6, -0.2, 0.025396825396, -0.001785714285]


# ------------------------------------------------------------------ stencil

def test_stencil_exact_on_polynomials():
    nxx, nzz, dh = 24, 28, 10.0
    x = (np.arange(nxx)[:, None] - 11.3) * dh
    z = (np.arange(nzz)[None, :] - 13.7) * dh
    inner = (slice(4, nxx - 4), slice(4, nzz - 4))
    for p in range(0, 10):
        for field, exact in ((x ** p + 0 * z, p * (p - 1) * x ** max(p - 2, 0) + 0 * z),
                             (0 * x + z ** p, 0 * x + p * (p - 1) * z ** max(p - 2, 0))):
            lap = ob.laplacian(field, dh)
            scale = np.abs(field[inner]).max() / dh ** 2 + 1.0
            err = np.abs(lap[inner] - exact[inner]).max() / scale
            assert err < 1e-10, (p, err)
    # degree 10 is no longer exact (the stencil is 8th order)
    f = x ** 10 + 0 * z
    lap = ob.laplacian(f, dh)
    exact = 90 * x ** 8 + 0 * z
    assert np.abs(lap[inner] - exact[inner]).max() > 1e-6 * np.abs(exact[inner]).max()


def test_stencil_weights_are_fornberg():
    # apply the stencil to a unit impulse: the response is the weight table
    nxx, nzz, dh = 21, 21, 1.0
    P = np.zeros((nxx, nzz))
    P[10, 10] = 1.0
    lap = ob.laplacian(P, dh)
    assert abs(lap[10, 10] - 2 * FORNBERG[0]) < 1e-11
    for m in range(1, 5):
        for d in ((m, 0), (-m, 0), (0, m), (0, -m)):
            assert abs(lap[10 + d[0], 10 + d[1]] - FORNBERG[m]) < 1e-11
    # no diagonal taps
    assert lap[11, 11] == 0.0 and lap[9, 12] == 0.0


def test_stencil_symbol_and_cfl():
    nxx, nzz, dh = 64, 12, 1.0
    j = np.arange(nxx)[:, None] + 0 * np.arange(nzz)[None, :]
    for kh in (0.3, 1.1, 2.0, math.pi):
        P = np.cos(kh * j).astype(np.float64)
        lap = ob.laplacian(P, dh)
        sym = FORNBERG[0] + 2 * sum(FORNBERG[m] * math.cos(m * kh) for m in range(1, 5))
        sel = (slice(4, nxx - 4), slice(4, nzz - 4))
        assert np.abs(lap[sel] - sym * P[sel]).max() < 1e-10
    # largest |symbol| (at kh = pi) fixes the 2-D leapfrog limit v dt/dh <= 0.5546
    smax = abs(FORNBERG[0] + 2 * sum(FORNBERG[m] * (-1) ** m for m in range(1, 5)))
    assert abs(smax - 6.5016) < 1e-3
    assert abs(2.0 / math.sqrt(2 * smax) - 0.5546) < 1e-4


# ------------------------------------------------------------------ forward

def small_case(nz=40, nx=50, nb=12, nt=300, dt=1e-3, dh=10.0, fmax=40.0):
    v = layered(nz, nx, [1500, 2200, 2800], anomaly=(4, 300.0), seed=3)
    vpad = ob.expand(v, nb)
    w, tlag = ob.wavelet(nt, dt, fmax)
    return v, vpad, w, tlag


def test_forward_linearity():
    v, vpad, w, tlag = small_case()
    nb = 12
    riz = np.array([14, 20, 30], np.int32)
    rix = np.array([15, 40, 60], np.int32)
    b = ob.abc_b1d(nb, 0.02)
    s1, _, _ = ob.forward(vpad, 10.0, 1e-3, 300, tlag, w, 16, 30, riz, rix, abc=True, b1d=b, nb=nb)
    s3, _, _ = ob.forward(vpad, 10.0, 1e-3, 300, tlag, 3.0 * w, 16, 30, riz, rix, abc=True, b1d=b,
                          nb=nb)
    assert rel_l2(s3, 3.0 * s1) < 1e-13


def test_forward_reciprocity():
    # swapping source and receiver scales the trace by (v_rec / v_src)^2
    v, vpad, w, tlag = small_case()
    A = (16, 25)
    B = (40, 52)
    sAB, _, _ = ob.forward(vpad, 10.0, 1e-3, 300, tlag, w, A[0], A[1], [B[0]], [B[1]])
    sBA, _, _ = ob.forward(vpad, 10.0, 1e-3, 300, tlag, w, B[0], B[1], [A[0]], [A[1]])
    vA = float(vpad[A[1], A[0]])
    vB = float(vpad[B[1], B[0]])
    assert abs(vA - vB) > 100.0            # a heterogeneous pair
    assert rel_l2(sAB[0] * vA ** 2, sBA[0] * vB ** 2) < 1e-11


def test_forward_zero_rim_and_sample0():
    v, vpad, w, tlag = small_case()
    seis, P, Q = ob.forward(vpad, 10.0, 1e-3, 300, tlag, w, 16, 30, [20], [30])
    for F in (P, Q):
        assert np.all(F[:4] == 0) and np.all(F[-4:] == 0)
        assert np.all(F[:, :4] == 0) and np.all(F[:, -4:] == 0)
        assert np.abs(F).max() > 0
    assert seis[0, 0] == 0.0


def test_sponge_absorbs():
    v, vpad, w, tlag = small_case(nt=900)
    nb = 12
    _, P0, _ = ob.forward(vpad, 10.0, 1e-3, 900, tlag, w, 16, 30, [20], [30])
    _, P1, _ = ob.forward(vpad, 10.0, 1e-3, 900, tlag, w, 16, 30, [20], [30], abc=True,
                          b1d=ob.abc_b1d(nb, 0.03), nb=nb)
    e0 = float(np.sum(P0 ** 2))
    e1 = float(np.sum(P1 ** 2))
    assert e1 < 0.05 * e0


def brute_force_forward(vpad, dh, dt, nt, tlag, w, s, rec, d=None):
    """Independent vectorised statement of compute_pressure / compute_seismogram
    (modeling.cu:301-353): inject, stencil on the current field, damp, sample."""
    nxx, nzz = vpad.shape
    dh, dt = float(np.float32(dh)), float(np.float32(dt))    # float parameters
    P = np.zeros((nxx, nzz))
    Q = np.zeros((nxx, nzz))
    seis = np.zeros((len(rec), nt))
    v2 = vpad.astype(np.float64) ** 2
    c = [2 * REF_WEIGHTS[0]] + REF_WEIGHTS[1:]
    for t in range(nt + tlag):
        if t < nt:
            P[s[1], s[0]] += w[t] / dh ** 2
        lap = np.zeros_like(P)
        I = (slice(4, nxx - 4), slice(4, nzz - 4))
        lap[I] = c[0] * P[I]
        for m in range(1, 5):
            lap[I] += c[m] * (P[4 + m:nxx - 4 + m, 4:-4] + P[4 - m:nxx - 4 - m, 4:-4]
                              + P[4:-4, 4 + m:nzz - 4 + m] + P[4:-4, 4 - m:nzz - 4 - m])
        new = Q.copy()
        new[I] = dt * dt * v2[I] * lap[I] / dh ** 2 + 2 * P[I] - Q[I]
        if d is not None:
            P[I] *= d[I]
            new[I] *= d[I]
        if t > tlag:
            for r, (iz, ix) in enumerate(rec):
                seis[r, t - tlag] = P[ix, iz]
        P, Q = new, P
    return seis, P, Q


def test_forward_against_brute_force_with_sponge():
    nz, nx, nb, nt = 20, 22, 8, 120
    v = layered(nz, nx, [1800, 2600], anomaly=(3, 200.0), seed=9)
    vpad = ob.expand(v, nb)
    w, tlag = ob.wavelet(nt, 1e-3, 60.0)
    b = ob.abc_b1d(nb, 0.05)
    nxx, nzz = vpad.shape
    d = np.array([[ob.lib().or_damper(b, i, j, nzz, nxx, nb) for i in range(nzz)]
                  for j in range(nxx)])
    rec = [(10, 9), (12, 20), (25, 30)]
    riz = np.array([r[0] for r in rec], np.int32)
    rix = np.array([r[1] for r in rec], np.int32)
    s_ref, P_ref, Q_ref = brute_force_forward(vpad, 10.0, 1e-3, nt, tlag, w, (11, 15), rec, d)
    s, P, Q = ob.forward(vpad, 10.0, 1e-3, nt, tlag, w, 11, 15, riz, rix, abc=True, b1d=b, nb=nb)
    assert rel_l2(s, s_ref) < 1e-12
    assert rel_l2(P, P_ref) < 1e-12 and rel_l2(Q, Q_ref) < 1e-12


# ------------------------------------------------------------------ reverse

def test_reverse_reconstructs_forward_fields():
    # after the source stops, the reverse recurrence retraces the forward
    # fields exactly: illum of K reverse steps = sum of u_{T-2-k}^2
    v, vpad, w, tlag = small_case(nt=200)
    nt, K = 200, 6
    T = nt + tlag
    _, PT, QT = ob.forward(vpad, 10.0, 1e-3, nt, tlag, w, 16, 30, [20], [30])
    fwd = {}
    for t in range(T - 2 - K + 1, T - 1):
        _, Pt, _ = ob.forward(vpad, 10.0, 1e-3, nt, t - nt, w, 16, 30, [20], [30])
        fwd[t] = Pt
    # a reverse pass of exactly K steps with zero data: nt_r = 1, tlag_r = K - 1
    zero = np.zeros((1, 1))
    _, ill = ob.reverse(vpad, 10.0, 1e-3, 1, K - 1, PT, QT, [20], [30], zero, 0)
    expect = sum(fwd[T - 2 - k] ** 2 for k in range(K))
    inner = (slice(4, -4), slice(4, -4))
    assert rel_l2(ill[inner], expect[inner]) < 1e-10


def test_reverse_image_and_gradient_definitions():
    # one reverse step by hand: S' = u_{T-1}, S = u_T, R_0 = idh2 d[nt-1]
    v, vpad, w, tlag = small_case(nt=60)
    dh, dt = 10.0, 1e-3
    rng = np.random.default_rng(0)
    S = rng.standard_normal(vpad.shape)
    Sp = rng.standard_normal(vpad.shape)
    for A in (S, Sp):
        A[:4] = A[-4:] = 0
        A[:, :4] = A[:, -4:] = 0
    data = np.zeros((1, 1))
    data[0, 0] = 2.5
    img, ill = ob.reverse(vpad, dh, dt, 1, 0, S, Sp, [20], [30], data, 0)
    grd, _ = ob.reverse(vpad, dh, dt, 1, 0, S, Sp, [20], [30], data, 1)
    lapS = ob.laplacian(Sp, dh)
    v2 = vpad.astype(np.float64) ** 2
    dt = float(np.float32(dt))
    Snew = dt * dt * v2 * lapS + 2 * Sp - S
    R = np.zeros_like(S)
    R[30, 20] = 2.5 / dh ** 2
    inner = (slice(4, -4), slice(4, -4))
    assert rel_l2(ill[inner], (Snew ** 2)[inner]) < 1e-13
    assert abs(img[30, 20] - Snew[30, 20] * R[30, 20]) < 1e-12 * abs(img[30, 20])
    assert np.count_nonzero(img) == 1
    g_expect = dt * R * lapS * v2
    assert abs(grd[30, 20] - g_expect[30, 20]) < 1e-12 * abs(g_expect[30, 20])


# ------------------------------------------------------------------ wavelet

def test_dft_pair():
    rng = np.random.default_rng(1)
    for n in (15, 16, 101):
        x = rng.standard_normal(n)
        re = np.zeros(n // 2 + 1)
        im = np.zeros(n // 2 + 1)
        ob.lib().or_dft_r2c(n, x, re, im)
        X = np.fft.rfft(x)
        assert np.abs(re - X.real).max() < 1e-10 and np.abs(im - X.imag).max() < 1e-10
        y = np.zeros(n)
        ob.lib().or_dft_c2r(n, re, im, y)
        assert np.abs(y / n - x).max() < 1e-12


def test_dft_matches_a_library_fft_at_wavelet_lengths():
    # the unnormalised r2c / c2r pair of src/modeling/modeling.cu:54-78 at the wavelet
    # lengths used (odd and even), against pocketfft (numpy); the c2r input
    # has arbitrary imaginary parts at DC and, for even n, at Nyquist, which
    # a Hermitian c2r transform ignores
    rng = np.random.default_rng(3)
    for n in (2000, 2001, 4149):
        x = rng.standard_normal(n)
        re, im = np.zeros(n // 2 + 1), np.zeros(n // 2 + 1)
        ob.lib().or_dft_r2c(n, x, re, im)
        X = np.fft.rfft(x)
        assert np.abs(re + 1j * im - X).max() < 1e-13 * np.abs(X).max()
        re, im = rng.standard_normal(n // 2 + 1), rng.standard_normal(n // 2 + 1)
        y = np.zeros(n)
        ob.lib().or_dft_c2r(n, re, im, y)
        ref = np.fft.irfft(re + 1j * im, n) * n
        assert np.abs(y - ref).max() < 1e-13 * np.abs(ref).max()


def ricker(nt, dt, fmax):
    dt, fmax = float(np.float32(dt)), float(np.float32(fmax))
    t0 = 2 * math.sqrt(math.pi) / fmax
    fc = fmax / (3 * math.sqrt(math.pi))
    td = np.arange(nt) * dt - t0
    a = math.pi ** 3 * fc * fc * td * td
    return 1e5 * (1 - 2 * a) * np.exp(-a)


@pytest.mark.parametrize("nt,dt,fmax", [(2001, 1e-3, 45.0), (2000, 1e-3, 50.0), (801, 6e-4, 36.0)])
def test_wavelet_is_half_derivative_of_ricker(nt, dt, fmax):
    w, tlag = ob.wavelet(nt, dt, fmax)
    r = ricker(nt, dt, fmax)
    W = np.fft.rfft(w)
    R = np.fft.rfft(r)
    f = np.fft.rfftfreq(nt, float(np.float32(dt)))
    H = np.sqrt(2 * np.pi * f) * np.exp(1j * np.pi / 4)
    if nt % 2 == 0:          # the Nyquist bin of a c2r transform keeps only its real part
        H[-1] = (R[-1] * H[-1]).real / R[-1] if R[-1] != 0 else 0
    sel = np.abs(R) > 1e-6 * np.abs(R).max()
    assert np.abs(W[sel] - H[sel] * R[sel]).max() < 1e-8 * np.abs(H * R).max()
    assert abs(W[0]) < 1e-6 * np.abs(W).max()          # DC removed
    t0 = 2 * math.sqrt(math.pi) / float(np.float32(fmax))
    assert tlag == int(np.float32(2.0) * np.sqrt(np.float32(math.pi)) / np.float32(fmax)
                       / np.float32(dt)) + 1
    assert abs(tlag - (t0 / dt + 1)) < 1.0 + 1e-9


# ------------------------------------------------------------------ sponge, pad, index

def test_cerjan_table_and_corners():
    nb, f = 17, 0.011
    b = ob.abc_b1d(nb, f)
    for k in range(nb):
        assert abs(b[k] - math.exp(-(f * (nb - k)) ** 2)) < 1e-15
    nzz, nxx = 60, 70
    for (i, j) in [(30, 35), (3, 35), (58, 35), (30, 2), (30, 69)]:
        d = ob.lib().or_damper(b, i, j, nzz, nxx, nb)
        inz, inx = nb <= i < nzz - nb, nb <= j < nxx - nb
        a = i if i < nb else nzz - 1 - i
        c = j if j < nb else nxx - 1 - j
        exp = 1.0 if inz and inx else (b[a] if inx else b[c])
        assert d == exp
    for (i, j) in [(0, 0), (5, 66), (57, 3), (59, 69)]:
        a = i if i < nb else nzz - 1 - i
        c = j if j < nb else nxx - 1 - j
        assert ob.lib().or_damper(b, i, j, nzz, nxx, nb) == b[a] + b[c] - 1.0
    # symmetric under z and x mirroring
    for (i, j) in [(2, 7), (10, 3)]:
        d = ob.lib().or_damper(b, i, j, nzz, nxx, nb)
        assert d == ob.lib().or_damper(b, nzz - 1 - i, j, nzz, nxx, nb)
        assert d == ob.lib().or_damper(b, i, nxx - 1 - j, nzz, nxx, nb)


def test_expand_reduce():
    rng = np.random.default_rng(2)
    v = rng.uniform(1500, 4000, (13, 9)).astype(np.float32)
    for nb in (1, 4, 7):
        e = ob.expand(v, nb)
        assert np.array_equal(e, np.pad(v, nb, mode="edge"))
        assert np.array_equal(ob.reduce(e, nb), v)


def test_grid_index_rounding():
    assert ob.grid_index(12.5, 25.0, 51) == 52           # half cell rounds up
    assert ob.grid_index(12.49, 25.0, 51) == 51
    assert ob.grid_index(0.0, 25.0, 51) == 51
    assert ob.grid_index(5000.0, 25.0, 51) == 200 + 51
    assert ob.grid_index(10.0, 10.0, 100) == 101


# ------------------------------------------------------------------ post-processing, step

def test_image_post_closed_form():
    nz, nx, dh = 30, 4, 5.0
    z = np.arange(nz) * dh
    illum = np.ones((nx, nz)) * 2.0
    image = np.tile(3.0 * z ** 2, (nx, 1)) * 2.0      # q = 3 z^2 -> -q'' = -6
    out = ob.image_post(image, illum, dh)
    assert np.all(out[:, 0] == 0) and np.all(out[:, -1] == 0)
    assert np.abs(out[:, 1:-1] + 6.0).max() < 1e-9


def test_gradient_post_mask_and_normalisation():
    nz, nx, dh, zmask = 40, 3, 10.0, 100.0
    z = np.arange(nz) * dh
    rng = np.random.default_rng(4)
    grad = rng.standard_normal((nx, nz)) * 1e-3 + np.tile(z ** 3, (nx, 1))
    illum = rng.uniform(1.0, 2.0, (nx, nz))
    out = ob.gradient_post(grad, illum, dh, zmask)
    izm = int(zmask / dh)
    assert np.all(out[:, :izm + 1] == 0) and np.all(out[:, -1] == 0)
    assert abs(np.abs(out).max() - 1.0) < 1e-15
    p = grad / illum
    raw = -(p[:, izm:-2] - 2 * p[:, izm + 1:-1] + p[:, izm + 2:]) / dh ** 2
    scale = np.abs(raw).max()
    assert np.abs(out[:, izm + 1:-1] - raw / scale).max() < 1e-12


def test_parabola_vertex():
    astar, c = 0.37, 2.0
    a = np.array([0.0, 0.2, 0.5])
    f = (a - astar) ** 2 * 4.0 + c
    assert abs(ob.parabola_step(a, f) - astar) < 1e-12
    # the reference's formula is the vertex of the interpolating parabola
    a = np.array([0.0, 0.02, 0.05]) * 3e-7
    f = np.array([10.0, 8.0, 7.5])
    s = ob.parabola_step(a, f)
    coef = np.polyfit(a, f, 2)
    assert abs(s - (-coef[1] / (2 * coef[0]))) < 1e-9 * abs(s)


# ------------------------------------------------------------------ random boundary

M64 = (1 << 64) - 1
G = 0x9E3779B97F4A7C15


def mix64(z):
    z ^= z >> 30
    z = (z * 0xBF58476D1CE4E5B9) & M64
    z ^= z >> 27
    z = (z * 0x94D049BB133111EB) & M64
    z ^= z >> 31
    return z


def draw(seed, stream, cell, k):
    key = mix64(seed ^ mix64((stream + G) & M64))
    h = mix64((key + (2 * cell + k + 1) * G) & M64)
    s = (h >> 40) - (1 << 23)
    return np.float32(s) * np.float32(1.0 / 8388608.0)


def rb_statement(vpad, nb, amp_abs, amp_rel, vlo, vhi, seed, stream):
    """Independent Python statement of the seeded random-boundary generator."""
    nxx, nzz = vpad.shape
    out = vpad.copy()
    f32 = np.float32
    for j in range(nxx):
        for i in range(nzz):
            xpad = j < nb or j >= nxx - nb
            zpad = i < nb or i >= nzz - nb
            if not (xpad or zpad):
                continue
            base = vpad[j, i]
            A = f32(amp_abs) + f32(f32(amp_rel) * base)
            val = base
            cell = i + j * nzz
            if xpad:
                dj = j if j < nb else nxx - 1 - j
                f = f32(1.0) - f32(f32(dj) / f32(nb))
                val = f32(val + f32(f32(f * A) * draw(seed, stream, cell, 0)))
            if zpad:
                di = i if i < nb else nzz - 1 - i
                f = f32(1.0) - f32(f32(di) / f32(nb))
                val = f32(val + f32(f32(f * A) * draw(seed, stream, cell, 1)))
            out[j, i] = min(max(val, f32(vlo)), f32(vhi))
    return out


def test_rb_generator_bit_exact_against_statement():
    v = layered(10, 12, [1500, 2500])
    nb = 6
    vpad = ob.expand(v, nb)
    for seed, stream, ab, rel in [(1, 0, 1000.0, 0.0), (2024, (3 << 32) | 17, 0.0, 0.4)]:
        vlo, vhi = np.float32(900.0), np.float32(3600.0)
        got = ob.rb_velocity(vpad, nb, ab, rel, vlo, vhi, seed, stream)
        exp = rb_statement(vpad, nb, ab, rel, vlo, vhi, seed, stream)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))


def test_rb_generator_properties():
    v = layered(60, 80, [1500, 2000, 3000])
    nb = 30
    vpad = ob.expand(v, nb)
    vlo, vhi = np.float32(800.0), np.float32(4500.0)
    out = ob.rb_velocity(vpad, nb, 0.0, 0.4, vlo, vhi, 77, 5)
    # pad only
    assert np.array_equal(out[nb:-nb, nb:-nb], v)
    # bounded by the ramp: |dv| <= (fx + fz) * 0.4 * base, clamped to [vlo, vhi]
    nxx, nzz = vpad.shape
    jj, ii = np.meshgrid(np.arange(nxx), np.arange(nzz), indexing="ij")
    dj = np.where(jj < nb, jj, np.where(jj >= nxx - nb, nxx - 1 - jj, nb))
    di = np.where(ii < nb, ii, np.where(ii >= nzz - nb, nzz - 1 - ii, nb))
    fx = np.where(dj < nb, 1 - dj / nb, 0.0)
    fz = np.where(di < nb, 1 - di / nb, 0.0)
    bound = (fx + fz) * 0.4 * vpad + 1e-3
    assert np.all(np.abs(out - vpad) <= bound)
    assert out.min() >= vlo and out.max() <= vhi
    # the outer columns are perturbed with the full amplitude, uniformly
    u = (out[0, nb:-nb] - vpad[0, nb:-nb]) / (0.4 * vpad[0, nb:-nb])
    assert abs(u.mean()) < 0.15 and 0.45 < u.std() < 0.7 and np.abs(u).max() <= 1.0
    # deterministic in the seed, different across streams
    assert np.array_equal(out, ob.rb_velocity(vpad, nb, 0.0, 0.4, vlo, vhi, 77, 5))
    assert not np.array_equal(out, ob.rb_velocity(vpad, nb, 0.0, 0.4, vlo, vhi, 77, 6))
    assert not np.array_equal(out, ob.rb_velocity(vpad, nb, 0.0, 0.4, vlo, vhi, 78, 5))


def test_rb_uniform_distribution():
    us = np.array([ob.lib().or_rb_uniform(11, 3, c, 0) for c in range(20000)])
    assert us.min() >= -1.0 and us.max() < 1.0
    hist, _ = np.histogram(us, bins=10, range=(-1, 1))
    assert np.all(np.abs(hist - 2000) < 200)
    assert abs(np.corrcoef(us[:-1], us[1:])[0, 1]) < 0.03
