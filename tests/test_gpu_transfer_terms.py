"""GPU: the transfer kernel's own per-point terms against the oracle.

k_replay (the transfer kernel of the split tracer, the cache replay and the
batch) evaluates density, source factor and ln z with exp_tab and log_tab,
sigma with the exact-uniform lookup where the table allows it (else the
bracketing one) and exp_tab, and the emission with exp_tab.  hadros_eval_transfer_point runs those
same device functions at given points; here every point's terms are compared
with a one-point ray of the oracle's integrate_kerr_ray_spectral
(src/radiative_transfer.cpp:640-685), whose per-bin tau is that point's dtau
and whose per-bin intensity is its emission term before absorption:

- FP64: ok flag equal, ln z within 2 ulp of numpy's, rho and every dtau_b
  and emission term within 1e-12 relative (below 1e-290 where the oracle's term is 0 or subnormal);
- FP32K (log-form emission, float sigma table): dtau_b and e^L within 1e-4
  relative;
- over all 8 profiles, 5 sources and 3 spectra, a log-uniform table (exact
  lookup), an uneven table (bracketing lookup) and energies at the table's
  edges.
"""
import math
import zlib

import numpy as np
import pytest

from hadros_amd import abi, ffi
from oracle import py_oracle as po

from image_metrics import rel
from scenes import SIG_E, SIG_S, parity_scenes

pytestmark = pytest.mark.gpu

A_BH = 0.95


def tables():
    rng = np.random.default_rng(7)
    uneven = np.sort(10 ** rng.uniform(4, 12, 40))
    uneven_s = 5.53e-36 * uneven ** 0.363 * (1 + 0.3 * rng.uniform(size=40))
    return [("log-uniform", SIG_E, SIG_S), ("uneven", uneven, uneven_s)]


def points(n, seed):
    rng = np.random.default_rng(seed)
    r = rng.uniform(1.0, 70.0, n)          # includes r < r_h + 1e-3 (r_h = 1.312 at a = 0.95)
    th = rng.uniform(0.0, math.pi, n)
    dl = 10 ** rng.uniform(-4, 0.5, n)
    z = 10 ** rng.uniform(-0.6, 0.6, n)    # E_mid z crosses the table's edges for the outer bins
    return r, th, dl, z


def oracle_terms(scene, se, ss, r, th, dl, z):
    dt = np.zeros((r.size, abi.MAX_BINS))
    em = np.zeros((r.size, abi.MAX_BINS))
    for k in range(r.size):
        pt = np.array([[r[k], th[k], 0.0, 0.0, 0.0, dl[k], z[k]]])
        _, bt, _, bi = po.integrate_ray(scene, se, ss, A_BH, pt)
        dt[k], em[k] = bt, bi
    return dt, em


@pytest.mark.parametrize("tname,se,ss", tables(), ids=[t[0] for t in tables()])
@pytest.mark.parametrize("name,scene", parity_scenes())
def test_transfer_terms_fp64(name, scene, tname, se, ss):
    r, th, dl, z = points(400, zlib.crc32(name.encode()))
    g = ffi.eval_transfer_point(scene, se, ss, A_BH, abi.PREC_FP64, r, th, dl, z)
    r_h = po.lib().or_horizon(A_BH)
    assert np.array_equal(g["ok"], r >= r_h + 1e-3)
    ok = g["ok"]
    lz = np.log(z[ok])
    assert np.all(np.abs(g["lnz"][ok] - lz) <= 2 * np.spacing(np.abs(lz)))
    rho_ref = np.array([po.lib().or_rho(scene.medium, r[k], th[k]) for k in range(r.size)])
    assert rel(g["rho"][ok], rho_ref[ok]).max() <= 1e-12
    dt, em = oracle_terms(scene, se, ss, r, th, dl, z)
    for got, ref, label in ((g["dtau"], dt, "dtau"), (g["emis"], em, "emission")):
        # relative error on normal terms; terms the oracle rounds to 0 or to a
        # subnormal (products of underflowing factors, formed in another
        # order) agree to below 1e-290 absolute
        big = ref > 1e-290
        assert np.all(np.abs(got[~big]) <= 1e-290), (label, int(np.sum(np.abs(got[~big]) > 1e-290)))
        err = rel(got[big], ref[big]) if big.any() else np.zeros(1)
        print(f"{name} {tname} {label}: {int(big.sum())} terms, max rel {err.max():.2e}")
        assert err.max() <= 1e-12, label


@pytest.mark.parametrize("tname,se,ss", tables(), ids=[t[0] for t in tables()])
@pytest.mark.parametrize("name,scene", parity_scenes()[:4])
def test_transfer_terms_fp32k(name, scene, tname, se, ss):
    r, th, dl, z = points(300, 17 + zlib.crc32(name.encode()))
    g = ffi.eval_transfer_point(scene, se, ss, A_BH, abi.PREC_FP32K, r, th, dl, z)
    dt, em = oracle_terms(scene, se, ss, r, th, dl, z)
    nb = po.bins(scene, se)[0].size          # the log-form columns past the scene's bins hold 0
    dt, em, gdt, gem = dt[:, :nb], em[:, :nb], g["dtau"][:, :nb], g["emis"][:, :nb]
    nz = dt > 0.0
    assert np.all(gdt[~nz] == 0.0)
    assert rel(gdt[nz], dt[nz]).max() <= 1e-4
    e = em > 1e-290
    assert np.all(np.exp(gem[~e]) <= 1e-290)
    if e.any():
        err = rel(np.exp(gem[e]), em[e])
        print(f"{name} {tname}: dtau max rel {rel(gdt[nz], dt[nz]).max():.2e}, emission max rel {err.max():.2e}")
        assert err.max() <= 1e-4


def test_transfer_terms_table_edges():
    """Bin energies 1e-14 relative inside and outside the table's first and
    last node: inside they count, outside dtau is 0 (the reference's range
    test), in both lookups.  (The kernels form E = E_mid z where the
    reference forms E_mid / (1 / z); the two differ by at most 2 ulp, so only
    energies within a few ulp of a node could be classified differently.)"""
    name, scene = parity_scenes()[1]
    for _, se, ss in tables():
        e0 = po.bins(scene, se)[0][0]
        zs = np.array([se[0] / e0 * (1 + 1e-14), se[0] / e0 * (1 - 1e-14),
                       se[-1] / e0 * (1 - 1e-14), se[-1] / e0 * (1 + 1e-14)])
        n = zs.size
        r, th, dl = np.full(n, 20.0), np.full(n, 1.5), np.full(n, 0.1)
        g = ffi.eval_transfer_point(scene, se, ss, A_BH, abi.PREC_FP64, r, th, dl, zs)
        assert np.all(g["ok"]) and np.all(g["rho"] > 0)
        dt, _ = oracle_terms(scene, se, ss, r, th, dl, zs)
        assert np.array_equal(dt[:, 0] > 0, [True, False, True, False]), dt[:, 0]
        assert np.array_equal(g["dtau"][:, 0] > 0, dt[:, 0] > 0), g["dtau"][:, 0]
        m = dt[:, 0] > 0
        assert rel(g["dtau"][m, 0], dt[m, 0]).max() <= 1e-12
