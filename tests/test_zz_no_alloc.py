"""The allocation-free get_greatestCluster entry point (sofia_cluster_greatest_noalloc)
against the allocating one (sofia_cluster_greatest) on the same engine state: the same
integers on two calls in a row, the second with a smaller system, through the C ABI with
device arrays the caller owns."""
import ctypes as C

import numpy as np
import pytest

from molecular_simulation_code_amd import system as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _engine():
    from molecular_simulation_code_amd import _lib
    if _lib.device_count() < 1:
        pytest.skip("no HIP device")
    from molecular_simulation_code_amd import host
    return host


def _hip():
    """The HIP runtime the library itself links (loaded once per process), for the
    caller-owned device arrays of the test."""
    hip = C.CDLL("libamdhip64.so.7")
    for name in ("hipMalloc", "hipMemcpy", "hipFree", "hipDeviceSynchronize"):
        getattr(hip, name).restype = C.c_int
    return hip


def _to_device(hip, a):
    a = np.ascontiguousarray(a)
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 16))) == 0
    assert hip.hipMemcpy(p, C.c_void_p(a.ctypes.data), C.c_size_t(a.nbytes), 1) == 0
    return p


def _noalloc(x, gid, mol, nmol, box, cut, dmin):
    from molecular_simulation_code_amd._lib import check, lib
    hip = _hip()
    bufs = [_to_device(hip, x.astype(np.float64)), _to_device(hip, gid.astype(np.int32)),
            _to_device(hip, mol.astype(np.int32)), _to_device(hip, np.full(4, -7, np.int32))]
    try:
        check(lib().sofia_cluster_greatest_noalloc(len(x), bufs[0], bufs[1], bufs[2], nmol,
                                                   float(box), float(cut), int(dmin), bufs[3], None))
        assert hip.hipDeviceSynchronize() == 0
        out = np.empty(4, np.int32)
        assert hip.hipMemcpy(C.c_void_p(out.ctypes.data), bufs[3], C.c_size_t(16), 2) == 0
        return out
    finally:
        for b in bufs:
            hip.hipFree(b)


def test_noalloc_greatest_cluster_equals_the_allocating_entry_point():
    host = _engine()
    from molecular_simulation_code_amd._lib import SofiaError
    cases = [(S.make_system(220, 1000, seed=93), 20), (S.make_system(60, 700, seed=94), 10)]
    for call, (s, steps) in enumerate(cases):
        e = host.Engine([host.ReplicaSpec(s, temp=300.0, gamma_ps=1.0, seed=7 + call)],
                        precision=host.PREC_MIXED)
        try:
            e.init()
            e.run(steps)
            x, gid = e.device_view()
            mol = np.where(s.moltype[gid - 1] == 1, s.mol[gid - 1], -1).astype(np.int32)
            for cut, dmin in ((7.5, 4), (6.0, 2)):
                tot, core, surf = e.greatest_cluster(cut, dmin)
                out = _noalloc(x, gid, mol, s.n_gly, s.box, cut, dmin)
                say(f"[no-alloc] call {call} ({s.n_gly} molecules) cut {cut} densmin {dmin}: "
                    f"allocating {tot[0]} ({core[0]} + {surf[0]}), allocation-free {out.tolist()}")
                assert out[3] == 0
                assert (out[0], out[1], out[2]) == (tot[0], core[0], surf[0])
        finally:
            e.close()
    # more molecules than the fixed scratch holds: refused on the host, nothing launched
    from molecular_simulation_code_amd._lib import check, lib
    hip = _hip()
    out_dev = _to_device(hip, np.zeros(4, np.int32))
    try:
        with pytest.raises(SofiaError) as ei:
            check(lib().sofia_cluster_greatest_noalloc(0, None, None, None, 40000, 30.0, 7.0, 3,
                                                       out_dev, None))
        assert ei.value.code == -3 and "32768" in str(ei.value)
    finally:
        hip.hipFree(out_dev)
