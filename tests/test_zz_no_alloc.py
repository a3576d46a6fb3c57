"""mmffhip_minimize_nk_noalloc (scratch in a library-owned device array, no workspace) against
mmffhip_minimize_nk (caller workspace), called through the C ABI: identical positions,
energies, statuses and counters on two calls in a row, the second with a smaller batch, for
global-Hessian plans (full matrix and symmetric blocks: the scratch holds the per-workgroup
slots, and the grid is capped to the slots that fit) and for an LDS plan (the scratch holds
only the work queue). A plan that needs
more than the scratch for one slot per compute unit is refused with a size error before
anything is launched."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _both(wl, place, block):
    from mlxmolkit_archive_amd import _lib, engine
    from mlxmolkit_archive_amd.packing import lpt_order
    L = _lib.load()
    db = engine.device_batch(wl.params, wl.conf_counts)
    o = engine.nk_options(max_iters=2000, grad_tol=1e-4, hess_place=place, block=block, ls_grad=True)
    x = torch.from_numpy(wl.x).cuda()
    order = torch.from_numpy(lpt_order(db.host)).cuda()
    C_ = db.n_confs
    outs = []
    for noalloc in (False, True):
        xo = torch.full_like(x, float("nan"))
        e = torch.full((C_,), float("nan"), dtype=torch.float64, device="cuda")
        st = torch.full((C_,), -7, dtype=torch.int32, device="cuda")
        cnt = torch.full((C_, 3), -7, dtype=torch.int32, device="cuda")
        if noalloc:
            print(f"noalloc call: {C_} conformers, place {place}, block {block}", flush=True)
            _lib.check(L.mmffhip_minimize_nk_noalloc(C.byref(db.desc), C.byref(o), x.data_ptr(), xo.data_ptr(),
                                                      e.data_ptr(), st.data_ptr(), cnt.data_ptr(), order.data_ptr(),
                                                      _lib.stream_ptr()))
        else:
            nb = C.c_size_t()
            _lib.check(L.mmffhip_minimize_nk_workspace(C.byref(db.desc), C.byref(o), C.byref(nb)))
            ws = torch.empty(int(nb.value), dtype=torch.uint8, device="cuda")
            _lib.check(L.mmffhip_minimize_nk(C.byref(db.desc), C.byref(o), x.data_ptr(), xo.data_ptr(), e.data_ptr(),
                                             st.data_ptr(), cnt.data_ptr(), order.data_ptr(), ws.data_ptr(),
                                             int(nb.value), _lib.stream_ptr()))
        torch.cuda.synchronize()
        outs.append((xo.cpu().numpy(), e.cpu().numpy(), st.cpu().numpy(), cnt.cpu().numpy()))
    return outs


@pytest.mark.parametrize("place,block", [(2, 256), (1, 128), (4, 256)])
def test_noalloc_matches_workspace_entry_point_twice(place, block):
    """Two calls in a row through both entry points, the second batch smaller. Full-matrix
    global plan: small-generator molecules (n <= 150): slots of 92 KB, of which the 64 MiB
    array holds 726, so the 750-conformer batch's grid (three 256-thread workgroups per CU) is
    capped to 726 workgroups. Symmetric-block global plan (the default placement): the same
    batches, slots of 48 KB, the grid not capped. LDS plan: the same batches' molecules with
    n <= 96."""
    from mlxmolkit_archive_amd import workloads
    from test_gpu_configs import _select
    big = workloads.build("small", n_mols=150, n_confs=5)
    small = workloads.build("small", n_mols=9, n_confs=3, seed=20261299)
    for wl in (big, small):
        if place == 1:
            wl = _select(wl, [i for i, p in enumerate(wl.params) if 3 * p.n_atoms <= 96])
        (x0, e0, s0, c0), (x1, e1, s1, c1) = _both(wl, place, block)
        assert np.array_equal(x0, x1) and np.array_equal(e0, e1)
        assert np.array_equal(s0, s1) and np.array_equal(c0, c1)
        assert np.isfinite(e1).all() and (s1 == 0).mean() > 0.9


def test_noalloc_refuses_slots_beyond_its_scratch():
    """A batch of 300 conformers whose largest molecule is declared as 100 atoms: a full
    global Hessian slot is 360 KB, and 256 of them (one per CU) exceed the 64 MiB array, so
    the call is refused before any launch."""
    from mlxmolkit_archive_amd import _lib, engine, workloads
    L = _lib.load()
    wl = workloads.build("small", n_mols=2, n_confs=150)
    db = engine.device_batch(wl.params, wl.conf_counts)
    desc = _lib.Batch.from_buffer_copy(db.desc)
    desc.max_atoms = 100
    o = engine.nk_options(max_iters=10, hess_place=_lib.HESS_GLOBAL, block=256)
    x = torch.from_numpy(wl.x).cuda()
    xo = torch.empty_like(x)
    e = torch.empty(db.n_confs, dtype=torch.float64, device="cuda")
    st = torch.empty(db.n_confs, dtype=torch.int32, device="cuda")
    n_cu = C.c_int32()
    _lib.check(L.mmffhip_device_info(C.byref(n_cu), None, None))
    assert n_cu.value * 360 * 1024 > (64 << 20)
    rc = L.mmffhip_minimize_nk_noalloc(C.byref(desc), C.byref(o), x.data_ptr(), xo.data_ptr(), e.data_ptr(),
                                       st.data_ptr(), None, None, _lib.stream_ptr())
    assert rc == -3
    assert b"64 MiB" in L.mmffhip_last_error()
