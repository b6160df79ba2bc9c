"""Host side of the Python binding (no GPU): the batch descriptors built
column-wise from slice offsets equal the ones built one by one, and result
fields are read from the C struct on demand."""
import ctypes as C

import numpy as np

import phj_gpu as G


def _parent(rows, keys, cols):
    r = G._Relation()
    r.rows = rows
    r.keys = keys
    r.n_cols = len(cols)
    r.on_device = 1
    for i, (t, ptr) in enumerate(cols):
        r.cols[i].type = t
        r.cols[i].data = ptr
    return G.DeviceRelation(None, r)


def _pairs(n):
    build = _parent(10**7, 0x1000, [(G.U32, 0x9000000), (G.U128, 0x40000000)])
    probe = _parent(8 * 10**7, 0x2000000, [(G.U8, 0x7000000), (G.U64, 0x5000000)])
    rng = np.random.default_rng(3)
    bo = np.concatenate([[0], np.cumsum(rng.integers(0, 2000, n))])
    po = np.concatenate([[0], np.cumsum(rng.integers(0, 9000, n))])
    pairs = [(build.slice(int(bo[i]), int(bo[i + 1])), probe.slice(int(po[i]), int(po[i + 1]))) for i in range(n)]
    # a view of a view keeps the root's offsets
    pairs[5] = (pairs[5][0].slice(3, 7), pairs[5][1])
    cfgs = [G.RadixConfig.for_build_rows(b.rows) for b, _ in pairs]
    return pairs, cfgs


def _bytes(descs):
    return C.string_at(descs, C.sizeof(descs))


def test_batch_descs_columnwise_equal_one_by_one():
    pairs, cfgs = _pairs(300)
    fast = G._batch_descs(pairs, cfgs, {"probe_chunk_rows": 4096})
    for b, _ in pairs:
        b._root = None  # forces the one-by-one path
    slow = G._batch_descs(pairs, cfgs, {"probe_chunk_rows": 4096})
    assert _bytes(fast) == _bytes(slow)
    d = fast[5]
    assert d.build.rows == 4 and d.build.keys == pairs[5][0]._rel.keys
    assert d.build.cols[1].data == pairs[5][0]._rel.cols[1].data
    assert d.tuning.probe_chunk_rows == 4096 and d.scheme == G.SCHEME_PHJ_PURE
    assert list(d.pass_bits[:d.n_passes]) == list(cfgs[5].pass_bits)


def test_result_fields_read_lazily():
    arr = (G._JoinResult * 3)()
    arr[1].e2e_wall_ms = 2.5
    arr[1].probe.wall_ms = 1.25
    arr[1].bep_evictions = 7
    r = G.JoinResult(None, arr[1])
    assert r.e2e_wall_ms == 2.5 and r.probe.wall_ms == 1.25 and r.bep_evictions == 7
    assert r.build_shuffle.wall_ms == 0.0
    try:
        r.no_such_field
    except AttributeError:
        pass
    else:
        raise AssertionError("unknown attribute must raise")
