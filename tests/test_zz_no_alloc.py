"""gm_mie_batch_noalloc (static device scratch, no allocation) against
gm_mie_batch through the C ABI: two calls in a row, the second on a smaller
batch, homogeneous (with a_n, b_n) and coated; the outputs must be
bitwise identical, since both run the same kernels on the same points (the
persistent grid size does not change any point's arithmetic)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(entry, x, y, m, m2, with_coeffs):
    import torch
    from geosmie_amd import _lib
    from geosmie_amd._lib import check, ptr
    L, h, s = _lib.lib(), _lib.ctx(), _lib.stream()
    n = x.size
    size = y if y is not None else x
    d_x = _lib.to_dev(x)
    d_y = None if y is None else _lib.to_dev(y)
    d_eps = _lib.cplx_dev(np.broadcast_to(np.asarray(m, np.complex128) ** 2, (n,)))
    d_eps2 = None if y is None else _lib.cplx_dev(
        np.broadcast_to(np.asarray(m2, np.complex128) ** 2, (n,)))
    d_props = torch.full((6 * n,), -1.0, dtype=torch.float64, device="cuda")
    d_nmax = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    d_off = d_an = d_bn = None
    if with_coeffs:
        d_off = torch.empty(n + 1, dtype=torch.int64, device="cuda")
        check(L.gm_coef_offsets(h, n, ptr(d_y if d_y is not None else d_x), ptr(d_off), s),
              "gm_coef_offsets")
        total = int(d_off[-1].item())
        d_an = torch.full((2 * total,), -1.0, dtype=torch.float64, device="cuda")
        d_bn = torch.full((2 * total,), -1.0, dtype=torch.float64, device="cuda")
    args = (h, n, ptr(d_x), ptr(d_y), ptr(d_eps), None, ptr(d_eps2), float(size.max()),
            ptr(d_off), ptr(d_an), ptr(d_bn), ptr(d_props), ptr(d_nmax), _lib.GM_MIE_SORT)
    if entry == "noalloc":
        d_status = torch.full((1,), 7, dtype=torch.int32, device="cuda")
        check(L.gm_mie_batch_noalloc(*args, ptr(d_status), s), "gm_mie_batch_noalloc")
        torch.cuda.synchronize()
        flags = int(d_status.item())
    else:
        check(L.gm_mie_batch(*args, s), "gm_mie_batch")
        flags = _lib.check_flags()
    out = {"props": d_props.cpu().numpy(), "nmax": d_nmax.cpu().numpy(), "flags": flags}
    if with_coeffs:
        out["an"], out["bn"] = d_an.cpu().numpy(), d_bn.cpu().numpy()
    return out


def _homog(n_r, n_l):
    r = np.logspace(np.log10(0.005), np.log10(20.0), n_r)
    lam = np.logspace(np.log10(0.3), np.log10(2.5), n_l)
    return (2 * np.pi * r[None, :] / lam[:, None]).ravel()


def _coated(n):
    rng = np.random.default_rng(20261202)
    y = 10 ** rng.uniform(np.log10(0.02), np.log10(209.0), n)
    ratio = rng.choice(np.linspace(0.3, 0.9, 7), n)
    gf = rng.choice(np.linspace(1.0, 2.5, 12), n)
    return ratio * y, y, 1.95 + 0.79j, 1.33 + (1.45 - 1.33) / gf ** 3


def _same(a, b):
    assert a["flags"] == 0 and b["flags"] == 0, (a["flags"], b["flags"])
    for k in a:
        if k != "flags":
            assert np.array_equal(a[k], b[k]), k


def test_noalloc_matches_batch_twice_homogeneous(gpu):
    big = _homog(2000, 10)                    # 20000 points, x up to 419
    small = _homog(300, 4)[::-1].copy()       # smaller batch, other order
    for x in (big, small):
        ref = _run("batch", x, None, 1.53 + 0.005j, None, True)
        got = _run("noalloc", x, None, 1.53 + 0.005j, None, True)
        _same(ref, got)
        assert np.all(got["props"][:x.size] > 0)


def test_noalloc_matches_batch_twice_coated(gpu):
    for n in (60000, 5000):
        x, y, m1, m2 = _coated(n)
        ref = _run("batch", x, y, m1, m2, n <= 5000)
        got = _run("noalloc", x, y, m1, m2, n <= 5000)
        _same(ref, got)


def test_noalloc_status_bits_and_limits(gpu):
    """A coated point with mu != 1 raises bit 1 in the caller's status (which
    the call first clears); a batch whose plan and one recurrence slot exceed
    the static array is refused before anything runs on it."""
    import torch
    from geosmie_amd import _lib
    from geosmie_amd._lib import ptr
    L, h, s = _lib.lib(), _lib.ctx(), _lib.stream()
    d_x = _lib.to_dev(np.array([1.0, 2.0]))
    d_y = _lib.to_dev(np.array([2.0, 3.0]))
    d_eps = _lib.cplx_dev(np.full(2, 2.25 + 0j))
    d_eps2 = _lib.cplx_dev(np.full(2, 1.44 + 0j))
    d_mu = _lib.cplx_dev(np.full(2, 1.5 + 0.6j))
    d_props = torch.empty(12, dtype=torch.float64, device="cuda")
    d_status = torch.full((1,), 1, dtype=torch.int32, device="cuda")
    rc = L.gm_mie_batch_noalloc(h, 2, ptr(d_x), ptr(d_y), ptr(d_eps), ptr(d_mu), ptr(d_eps2), 3.0,
                                None, None, None, ptr(d_props), None, _lib.GM_MIE_SORT,
                                ptr(d_status), s)
    assert rc == 0
    torch.cuda.synchronize()
    assert int(d_status.item()) == 2
    # n_max(1e5) rows x 64 lanes x 88 bytes is far beyond the static array
    rc = L.gm_mie_batch_noalloc(h, 2, ptr(d_x), ptr(d_y), ptr(d_eps), None, ptr(d_eps2), 1e5,
                                None, None, None, ptr(d_props), None, _lib.GM_MIE_SORT,
                                ptr(d_status), s)
    assert rc == -1
    assert b"static scratch holds" in L.gm_last_error()
    # and n_max(1e6) beyond what one wavefront's slot may span at all
    rc = L.gm_mie_batch_noalloc(h, 2, ptr(d_x), ptr(d_y), ptr(d_eps), None, ptr(d_eps2), 1e6,
                                None, None, None, ptr(d_props), None, _lib.GM_MIE_SORT,
                                ptr(d_status), s)
    assert rc == -1
    assert b"too large" in L.gm_last_error()
    rc = L.gm_mie_batch_noalloc(h, 2, ptr(d_x), ptr(d_y), ptr(d_eps), None, ptr(d_eps2), 3.0,
                                None, None, None, ptr(d_props), None, _lib.GM_MIE_SORT, None, s)
    assert rc == -1
    torch.cuda.synchronize()
