"""Pins of the oracle's 2D/3D step composition, independent of the oracle.

Each check recomputes a part of one split step with numpy/scipy from the
mathematics the reference implements, and compares it with the FP64 oracle
(oracle/fgpe_oracle.c):

  * the y and z FFT axes: or_fft_forward / or_fft_backward against
    numpy.fft.fftn / ifftn on grids with three different axis lengths
    (SRC/cgpe3d.f90:1748-1784, plans 852-853);
  * the linear part FFT -> KE -> SOC -> BFT against the per-wavevector
    product expm(-i CDT 2(gx kx Sx + gy ky Sy)) . expm(-i CDT (k^2 + 2 gz kz Sz))
    with the reference's positive-Nyquist wavenumbers (SRC/cgpe3d.f90:702-726,
    875-1004; the gz part sits in KE and is applied first, the two factors do
    not commute);
  * the whole loop body (SRC/cgpe3d.f90:452-495) as that numpy linear part
    followed by the separately pinned SE and SP and, in imaginary time, NORMT;
  * real-time conservation of sum |phi|^2 (every factor is unitary) and the
    monotone energy decrease of imaginary-time relaxation;
  * the SOC = 0 imaginary-time projection: the per-component factors
    sigma_m that one step applies, recovered from the step itself, satisfy
    sigma_1 sigma_3 = sigma_2^2 (spin 1, SRC/cgpe3d.f90:462-467) or
    log sigma_m affine in m (spin 2, the NEWTON ansatz exp((x1 + m x2) DT),
    SRC/cgpe3d.f90:468-481), on top of total norm 1 and magnetisation MAG.
"""
import numpy as np
import pytest

import oracle_ctypes as O
from cases import NA_SPIN2, case, noisy
from test_oracle_pins import SPIN_MATS


def kaxis(n, d):
    """Wavenumbers in FFT order with a positive Nyquist entry
    (SRC/cgpe3d.f90:702-726: k_i = (i-1) 2 pi / L for i <= N/2 + 1)."""
    k = 2 * np.pi * np.fft.fftfreq(n, d=d)
    k[n // 2] = np.pi / d
    return k


def spatial_axes(dim):
    return tuple(range(4 - dim, 4))


def kgrids(p, dim):
    ks = [kaxis(p.n[a], p.d[a]) if a < dim else np.zeros(1) for a in range(3)]
    return np.meshgrid(ks[2], ks[1], ks[0], indexing="ij")  # KZ, KY, KX


def herm_expm(g, cdt):
    """exp(-i cdt G) for a batch of Hermitian matrices G (..., ms, ms)."""
    w, v = np.linalg.eigh(g)
    return np.einsum("...ij,...j,...kj->...ik", v, np.exp(-1j * cdt * w), v.conj())


def numpy_linear_step(p, dim, spin, phi):
    """FFT, KE, SOC, BFT of one step from the mathematics (no oracle code)."""
    sx, sy, sz = SPIN_MATS[spin]
    cdt = -1j * p.dt if p.imag else p.dt + 0j
    KZ, KY, KX = kgrids(p, dim)
    ax = spatial_axes(dim)
    ph = np.fft.fftn(phi, axes=ax)
    gx, gy, gz = p.gamma[0], p.gamma[1], p.gamma[2]
    # KE: exp(-i CDT (k^2 + 2 gz kz Sz)), Sz diagonal (3D carries gz; 2D/1D gz = 0)
    szd = np.real(np.diag(sz))
    k2 = KX**2 + KY**2 + KZ**2
    for m in range(2 * spin + 1):
        ph[m] *= np.exp(-1j * cdt * (k2 + (2 * gz * szd[m] * KZ if dim == 3 else 0.0)))
    # SOC: exp(-i CDT 2 (gx kx Sx + gy ky Sy)) per (kx, ky)
    g = 2 * (gx * KX[..., None, None] * sx + gy * KY[..., None, None] * sy)
    u = herm_expm(g, cdt)
    ph = np.einsum("zyxab,bzyx->azyx", u, ph)
    return np.fft.ifftn(ph, axes=ax)


FFT_SHAPES = [(2, (24, 16, 1)), (2, (12, 32, 1)), (3, (16, 12, 8)), (3, (8, 6, 24))]


@pytest.mark.parametrize("dim,n", FFT_SHAPES)
def test_fft_axes_match_numpy_fftn(dim, n):
    """Forward: unnormalised exp(-2 pi i jk/n) along every axis; backward:
    exp(+2 pi i jk/n) divided by NX NY NZ (SRC/cgpe3d.f90:1748-1784)."""
    inp = case(dim, 1, n[:dim], (0.3,) * dim, 1e-3, (0.0,) * dim)
    p, _ = O.derive(inp)
    o = O.Oracle(p)
    rng = np.random.default_rng(sum(n))
    phi = rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)
    o.set_phi(phi)
    o.call("or_fft_forward")
    want = np.fft.fftn(phi, axes=spatial_axes(dim))
    assert np.abs(o.get_phif() - want).max() < 1e-12 * np.abs(want).max()
    f = rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)
    o.set_phif(f)
    o.call("or_fft_backward")
    want = np.fft.ifftn(f, axes=spatial_axes(dim))
    assert np.abs(o.get_phi() - want).max() < 1e-12 * np.abs(want).max()
    o.close()


LIN_CASES = [(dim, spin, imag) for dim in (1, 2, 3) for spin in (1, 2) for imag in (0, 1)]


def _lin_case(dim, spin, imag, dt=0.01):
    n = {1: (64,), 2: (24, 16), 3: (16, 12, 8)}[dim]
    gam = {1: (0.9,), 2: (0.7, 1.1), 3: (0.5, 0.8, 0.6)}[dim]
    inp = case(dim, spin, n, (0.3,) * dim, dt, gam, imag=imag, natoms=2000, nu=(20.0, 23.0, 27.0))
    return O.derive(inp)[0]


@pytest.mark.parametrize("dim,spin,imag", LIN_CASES)
def test_linear_step_is_per_k_exponential(dim, spin, imag):
    """FFT -> KE -> SOC -> BFT on white noise (which fills the Nyquist planes)
    equals the numpy per-wavevector exponential product.  A flipped gz sign,
    a negative Nyquist entry, the SOC factor applied before the gz factor or
    a transform along the wrong stride all fail here."""
    p = _lin_case(dim, spin, imag)
    o = O.Oracle(p)
    rng = np.random.default_rng(dim * 100 + spin * 10 + imag)
    phi = rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)
    o.set_phi(phi)
    for f in ("or_fft_forward", "or_ke", "or_soc", "or_fft_backward"):
        o.call(f)
    got = o.get_phi()
    o.close()
    want = numpy_linear_step(p, dim, spin, phi)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def test_linear_step_order_matters_in_3d():
    """The check above can tell the reference's order (gz factor first) from
    the reversed one: the two products differ by far more than its bar."""
    p = _lin_case(3, 2, 0)
    sx, sy, sz = SPIN_MATS[2]
    KZ, KY, KX = kgrids(p, 3)
    cdt = p.dt + 0j
    a = herm_expm(2 * (p.gamma[0] * KX[..., None, None] * sx + p.gamma[1] * KY[..., None, None] * sy), cdt)
    b = herm_expm(2 * p.gamma[2] * KZ[..., None, None] * sz, cdt)
    assert np.abs(a @ b - b @ a).max() > 1e-4


def _gaussian_state(p, dim, spin, seed):
    o = O.Oracle(p)
    o.initial_guess()
    o.normt()
    phi = noisy(o.get_phi(), seed=seed, amp=0.05)
    o.close()
    return phi


STEP_CASES = [(2, 1, 0), (2, 2, 0), (2, 2, 1), (3, 1, 1), (3, 2, 0), (3, 2, 1), (1, 1, 0), (1, 2, 1)]


@pytest.mark.parametrize("dim,spin,imag", STEP_CASES)
def test_whole_step_is_linear_part_then_se_sp(dim, spin, imag):
    """One loop body (SRC/cgpe3d.f90:452-495, SO coupling on) equals the numpy
    linear part followed by SE and SP (each pinned against expm and the
    interaction energy) and, in imaginary time, NORMT."""
    p = _lin_case(dim, spin, imag, dt=2e-3)
    phi = _gaussian_state(p, dim, spin, seed=dim + 3 * spin + 7 * imag)
    o = O.Oracle(p)
    o.set_phi(phi)
    o.run(1)
    got = o.get_phi()
    o.set_phi(numpy_linear_step(p, dim, spin, phi))
    o.call("or_se")
    o.call("or_sp")
    if imag:
        o.normt()
    want = o.get_phi()
    o.close()
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("dim,spin", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_real_time_steps_conserve_sum_of_squares(dim, spin):
    """Real time: FFT/BFT, KE, SOC, SE (exp of the Hermitian Hbar) and SP are
    all unitary, so sum |phi|^2 over grid and components is constant."""
    p = _lin_case(dim, spin, 0, dt=2e-3)
    phi = _gaussian_state(p, dim, spin, seed=5)
    o = O.Oracle(p)
    o.set_phi(phi)
    s0 = np.sum(np.abs(phi) ** 2)
    for _ in range(4):
        o.run(5)
        s = np.sum(np.abs(o.get_phi()) ** 2)
        assert abs(s - s0) < 1e-12 * s0
    o.close()


@pytest.mark.parametrize("dim,spin", [(2, 1), (2, 2), (3, 2)])
def test_imaginary_time_energy_decreases(dim, spin):
    """Imaginary time with SO coupling: the normalised flow lowers ENERGY
    monotonically from a noisy start."""
    n = {2: (32, 32), 3: (16, 16, 16)}[dim]
    gam = {2: (0.5, 0.4), 3: (0.3, 0.4, 0.2)}[dim]
    inp = case(dim, spin, n, (0.5,) * dim, 2e-3, gam, imag=1, natoms=2000, nu=(20.0, 23.0, 27.0))
    p, _ = O.derive(inp)
    phi = _gaussian_state(p, dim, spin, seed=9)
    o = O.Oracle(p)
    o.set_phi(phi)
    o.normt()
    ens = [o.energy()[0]]
    for _ in range(6):
        o.run(10)
        ens.append(o.energy()[0])
    o.close()
    assert all(b < a for a, b in zip(ens, ens[1:])), ens


def _projection_factors(p, dim, spin, phi):
    """Per-component factor one SOC = 0 imaginary-time step applies after SP:
    phi_after / phi_before_projection, constant over the grid."""
    o = O.Oracle(p)
    o.set_phi(phi)
    o.run(1)
    after = o.get_phi()
    pre_norms = o.last_norms()
    o.set_phi(phi)
    for f in ("or_fft_forward", "or_ke", "or_soc", "or_fft_backward", "or_se", "or_sp"):
        o.call(f)
    before = o.get_phi()
    o.close()
    ms = 2 * spin + 1
    r = np.empty(ms)
    for m in range(ms):
        big = np.abs(before[m]) > 1e-3 * np.abs(before[m]).max()
        ratio = after[m][big] / before[m][big]
        assert np.abs(ratio.imag).max() < 1e-13
        assert np.ptp(ratio.real) < 1e-12 * abs(ratio.real.mean())
        r[m] = ratio.real.mean()
    return r, pre_norms


@pytest.mark.parametrize("dim,mag", [(1, 0.3), (2, -0.2), (3, 0.25), (3, -0.5)])
def test_spin1_projection_factors(dim, mag):
    """sigma_2 of SRC/cgpe3d.f90:462-465 (3D; 1D/2D SRC/cgpe1d.f90:447-448) is
    the positive solution of sigma_1 sigma_3 = sigma_2^2 with the two
    constraints; NORMT rescales all three alike, so the product relation
    survives in the applied factors."""
    n = {1: (256,), 2: (64, 64), 3: (32, 32, 32)}[dim]
    d = {1: (0.05,), 2: (0.1, 0.1), 3: (0.4, 0.4, 0.4)}[dim]
    inp = case(dim, 1, n, d, 5e-4, (0.0,) * dim, mag=mag, option_fpc=1,
               natoms=100000 if dim == 3 else 10000, nu=(20.0, 20.0, 20.0) if dim == 3 else (20.0, 400.0, 400.0))
    p, _ = O.derive(inp)
    phi = _gaussian_state(p, dim, 1, seed=4)
    r, nrm = _projection_factors(p, dim, 1, phi)
    assert abs(r[0] * r[2] - r[1] ** 2) < 1e-12 * r[1] ** 2
    after = nrm * r**2
    assert abs(after.sum() - 1.0) < 1e-12
    assert abs(after[0] - after[2] - mag) < 1e-12


@pytest.mark.parametrize("dim,mag", [(1, 0.5), (2, 0.6), (2, -1.1), (3, 0.4)])
def test_spin2_projection_factors(dim, mag):
    """NEWTON (SRC/cgpe3d.f90:468-481, 1516-1588; 2D SRC/cgpe2d.f90:1286-1359)
    gives sigma_m = exp((x1 + m x2) DT): log sigma_m is affine in m, and the
    projected norms sum to 1 with magnetisation MAG (to the Newton tolerance)."""
    n = {1: (256,), 2: (64, 64), 3: (32, 32, 32)}[dim]
    d = {1: (0.05,), 2: (0.1, 0.1), 3: (0.4, 0.4, 0.4)}[dim]
    inp = case(dim, 2, n, d, 1.25e-4, (0.0,) * dim, mag=mag, option_fpc=1,
               natoms=100000 if dim == 3 else 10000, nu=(20.0, 20.0, 20.0) if dim == 3 else (20.0, 400.0, 400.0),
               species=NA_SPIN2 if dim == 1 else None)
    p, _ = O.derive(inp)
    phi = _gaussian_state(p, dim, 2, seed=6)
    r, nrm = _projection_factors(p, dim, 2, phi)
    lr = np.log(r)
    m = np.array([2.0, 1.0, 0.0, -1.0, -2.0])
    fit = np.polyfit(m, lr, 1)
    assert np.abs(np.polyval(fit, m) - lr).max() < 1e-12
    after = nrm * r**2
    assert abs(after.sum() - 1.0) < 1e-12
    assert abs(np.dot(m, after) - mag) < 1e-8
