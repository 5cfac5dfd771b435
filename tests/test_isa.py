"""Properties of the device code as built (gfx950 assembly of
emopt_amd/csrc/emfdtd.hip with the build's own flags), checked on the CPU:

* no kernel writes through the scalar data cache (scalar stores, scalar
  atomics, scalar cache write-back): every store is a vector store;
* every function addresses memory through the global (or constant) address
  space: no flat loads or stores, which would also count on the scalar-memory
  counter and make each wait for a descriptor load wait for the field loads in
  flight (the batch passes, whose array pointers come from member descriptors
  in memory, were the case that mattered).

This file names those instructions, so it stays on the CPU side."""
import os
import re
import subprocess

import pytest

from emopt_amd import build as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOT = ("k_pass_h", "k_pass_e", "k_batch_h", "k_batch_e", "k_batch_phasor")


@pytest.fixture(scope="module")
def isa(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("isa") / "emfdtd.s")
    flags = [f for f in B.FLAGS if f not in ("-shared", "-fPIC")]
    cmd = [B.hipcc(), *flags, "-DEMF_EXPORT=1", "-I", os.path.join(ROOT, "include"),
           "--cuda-device-only", "-S", os.path.join(ROOT, "emopt_amd", "csrc", "emfdtd.hip"),
           "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    with open(out) as f:
        return f.read().split("\n")


def _functions(lines):
    """(symbol, body lines) of every function in the assembly."""
    starts = [(i, l.split(":")[0]) for i, l in enumerate(lines)
              if re.match(r"^_Z\S+:", l) and "@" in l]
    for i, name in starts:
        j = i + 1
        while j < len(lines) and not lines[j].startswith(".Lfunc_end"):
            j += 1
        yield name, lines[i:j]


def test_no_scalar_cache_writes(isa):
    bad = [l.strip() for l in isa
           if re.match(r"^\s+(s_store|s_atomic|s_buffer_store|s_buffer_atomic|s_scratch_store|"
                       r"s_dcache_wb|s_dcache_discard)", l)]
    assert not bad, bad[:5]


def test_no_flat_accesses(isa):
    seen = set()
    for name, body in _functions(isa):
        seen.update(h for h in HOT if h in name)
        flat = [l.strip() for l in body if re.match(r"^\s+flat_(load|store|atomic)", l)]
        assert not flat, (name, flat[:3])
    assert seen == set(HOT), seen
