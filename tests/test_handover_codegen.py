"""CPU guard of the code the cross-XCD hand-over relies on (device.hpp).

The bottom-up kernels hand a node from one climbing thread to another,
possibly on another XCD, with relaxed agent-scope atomics and waits instead of
acquire/release (which would write back and invalidate a whole XCD L2 per
hand-over).  That is correct because of what the compiler emits for gfx950,
which this test pins by compiling the primitives with the library's flags and
reading the assembly:

- values another XCD may read are stored with the agent-scope bit (sc1), so
  they go through to the point of coherence shared by the XCDs;
- the arrival atomic (exchange or counter) is issued only after
  `s_waitcnt vmcnt(0)`: every earlier store of the wave has been acknowledged;
- the arrival's result is waited for (`vmcnt(0)`) before anything uses it, and
  the loads of the handed-over values are device-coherent (sc1) and issued
  after that wait -- also after a peek, whose value they depend on.

A compiler that emitted these differently would fail here before any GPU run.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import pytest

import atrbvh_amd.build as B

SNIPPET = r"""
#include "device.hpp"
using namespace atr;
extern "C" __global__ void k_publish(float4* box, uint64_t* slot, uint64_t v, uint64_t* out)
{
    st_coh(&box[threadIdx.x], make_float4(1.f, 2.f, 3.f, 0.f));
    out[threadIdx.x] = exchange_crossing(&slot[threadIdx.x], v);
}
extern "C" __global__ void k_take(const float4* box, uint64_t* slot, float4* out)
{
    const uint64_t o = peek_crossing(&slot[threadIdx.x]);
    if (o != ~0ull) out[threadIdx.x] = ld<true>(&box[(uint32_t)o]);
}
extern "C" __global__ void k_counter(uint32_t* c, uint32_t* out)
{
    out[threadIdx.x] = arrive_crossing(&c[threadIdx.x]);
}
"""


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    if not os.path.exists(B.HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("codegen")
    src = d / "handover.hip"
    src.write_text(SNIPPET)
    out = d / "handover.s"
    cmd = [B.HIPCC, *B.COMMON, *B.DEVICE, "--cuda-device-only", "-S", "-I", B.CSRC,
           "-x", "hip", str(src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    text = out.read_text()
    kernels = {}
    for name in ("k_publish", "k_take", "k_counter"):
        m = re.search(rf"^{name}:(.*?)s_endpgm", text, re.S | re.M)
        assert m, name
        kernels[name] = [ln.strip() for ln in m.group(1).splitlines()
                         if ln.strip() and not ln.strip().startswith(";")]
    return kernels


def _index(lines, pred, start=0):
    for i in range(start, len(lines)):
        if pred(lines[i]):
            return i
    raise AssertionError("pattern not found after line %d:\n%s" % (start, "\n".join(lines)))


def _is_full_wait(ln):
    return ln.startswith("s_waitcnt") and "vmcnt(0)" in ln


def test_coherent_store_then_wait_then_exchange(asm):
    k = asm["k_publish"]
    stores = [i for i, ln in enumerate(k) if ln.startswith("global_store")]
    assert len(stores) >= 2
    # the box stores carry the agent-scope bit
    for i in stores[:2]:
        assert re.search(r"\bsc1\b", k[i]), k[i]
    swap = _index(k, lambda ln: ln.startswith("global_atomic_swap_x2"))
    assert swap > stores[1]
    # every store before the exchange is acknowledged before it issues
    wait = max(i for i in range(stores[1], swap) if _is_full_wait(k[i]))
    assert stores[1] < wait < swap
    # the exchange returns its value and that value is waited for before use
    assert re.search(r"\bsc0\b", k[swap]), k[swap]
    _index(k, _is_full_wait, swap + 1)


def test_peek_then_dependent_coherent_loads(asm):
    k = asm["k_take"]
    peek = _index(k, lambda ln: ln.startswith("global_load_dwordx2"))
    assert re.search(r"\bsc1\b", k[peek]), k[peek]
    wait = _index(k, _is_full_wait, peek + 1)
    loads = [i for i in range(wait + 1, len(k)) if k[i].startswith("global_load")]
    assert len(loads) >= 2
    for i in loads:
        assert re.search(r"\bsc1\b", k[i]), k[i]


def test_counter_arrival_is_fenced_by_waits(asm):
    k = asm["k_counter"]
    add = _index(k, lambda ln: ln.startswith("global_atomic_add"))
    assert any(_is_full_wait(k[i]) for i in range(add))
    assert re.search(r"\bsc0\b", k[add]), k[add]
    _index(k, _is_full_wait, add + 1)
