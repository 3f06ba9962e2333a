"""The allocation-free coupling entry point (eq_batch_couple_traps_dev_noalloc)
against the allocating one (eq_batch_couple_traps_dev), through the C ABI:
two calls in a row, the second with fewer traps and cells, on two identical
batches; fields, per-cell samples and counts bit-identical (the same kernels
run on the same inputs).  A call whose scratch exceeds the fixed buffer is
refused with EQ_ERR_INVALID and leaves the fields as they were."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_coupling_matches_allocating_entry(eqf, oracle):
    import torch
    from eq_amd import workloads as W
    traps = W.config3_traps(512, 3)
    ids = [75, 475, 130, 11, 10, 0]
    sel = [traps[k] for k in ids]
    descs = [eqf.field_desc(t.nx, t.ny, t.width, t.height, 0.1, D, g, 0.0, W.TOP_DIRICHLET)
             for t in sel for D, g in zip(t.D, t.gamma)]
    rng = np.random.default_rng(12)
    u0 = [rng.random(t.nodes) for t in sel for _ in range(2)]
    dev = torch.device("cuda")
    A = eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(eqf.MIXED_GMG_CG))
    B = eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(eqf.MIXED_GMG_CG))
    try:
        for k, u in enumerate(u0):
            A.upload(k, u)
            B.upload(k, u)
        # call 1: every selected trap; call 2: traps 1..3 only, fewer cells
        for call, (t0, t1) in enumerate([(0, len(sel)), (1, 4)]):
            cells = np.concatenate([sel[t].cells for t in range(t0, t1)])
            q = np.concatenate([sel[t].rates for t in range(t0, t1)]) * 0.1
            off = np.concatenate([[0], np.cumsum([sel[t].cells.shape[0] for t in range(t0, t1)])]).astype(np.int64)
            cells_d = torch.from_numpy(cells).to(dev).contiguous()
            q_d = torch.from_numpy(q).to(dev).contiguous()
            sa = torch.zeros((int(off[-1]), 2), dtype=torch.float64, device=dev)
            sb = torch.zeros_like(sa)
            ca = torch.zeros(int(off[-1]), dtype=torch.int32, device=dev)
            cb = torch.zeros_like(ca)
            print(f"noalloc call {call}: traps {t0}..{t1 - 1}, {int(off[-1])} cells", flush=True)
            A.couple_traps_dev(t0, off, cells_d.data_ptr(), 10.0, q_d.data_ptr(), sa.data_ptr(), ca.data_ptr())
            B.couple_traps_dev_noalloc(t0, off, cells_d.data_ptr(), 10.0, q_d.data_ptr(), sb.data_ptr(),
                                       cb.data_ptr())
            A.sync()
            B.sync()
            assert torch.equal(ca, cb), call
            assert torch.equal(sa, sb), call
            for k in range(len(descs)):
                assert np.array_equal(A.download(k), B.download(k)), (call, k)
        # beyond the fixed buffer: refused, fields untouched
        before = [B.download(k) for k in range(2)]
        many = np.tile(sel[1].cells[:1], (800_000, 1))
        many_d = torch.from_numpy(many).to(dev).contiguous()
        qm = torch.ones((many.shape[0], 2), dtype=torch.float64, device=dev)
        with pytest.raises(eqf.EqError) as e:
            B.couple_traps_dev_noalloc(0, np.array([0, many.shape[0]], np.int64), many_d.data_ptr(), 10.0,
                                       qm.data_ptr())
        assert e.value.code == 1 and "fixed buffer" in str(e.value)
        B.sync()
        for k in range(2):
            assert np.array_equal(B.download(k), before[k])
    finally:
        A.close()
        B.close()
