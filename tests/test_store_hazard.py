"""The gfx950 assembly of the sweep kernels holds no MUBUF store of more than 64 bits
whose data VGPRs a VALU instruction overwrites within two wait states.

Such a store reads its data VGPRs after it issues.  The compiler separates it from a
later write of those VGPRs only when the store's soffset is not an SGPR; with an SGPR
soffset it scheduled the write directly after the store, and the store then sometimes
wrote the new register values (a few elements per launch, differing between runs).  The
kernels therefore issue their 16-byte stores with a zero soffset (buf_store_f4), and
this test compiles the per-size sweep kernels and the compiled mixed-radix plans to
assembly and checks every wide store in them."""
import os
import re
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "qstem_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "--cuda-device-only", "-S"]

WIDE_STORE = re.compile(r"(buffer_store_dwordx[34]|buffer_store_b96|buffer_store_b128)\s+(\S+),\s*(\S+),\s*(\S+),\s*(\S+)")


def _vregs(tok):
    m = re.match(r"v\[(\d+):(\d+)\]", tok)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"v(\d+)$", tok)
    return {int(m.group(1))} if m else set()


def hazards(asm):
    """(line, store, writer) for each wide MUBUF store whose data VGPRs a VALU
    instruction writes within two wait states (one per instruction, s_nop n: n + 1)."""
    lines = asm.split("\n")
    out = []
    for i, line in enumerate(lines):
        m = WIDE_STORE.match(line.strip())
        if not m:
            continue
        data = _vregs(m.group(2))
        waits = 0
        for nxt in lines[i + 1:i + 16]:
            t = nxt.strip()
            if not t or t.startswith(";") or (t.startswith(".") and not t.startswith(".LBB")):
                continue
            if t.endswith(":"):  # a branch target: the wait states are counted from here on another path
                break
            op = t.split()[0]
            if op == "s_nop":
                waits += int(t.split()[1], 0) + 1
            else:
                if op.startswith("v_") and not op.startswith(("v_cmp", "v_readlane", "v_readfirstlane")):
                    if _vregs(t.split()[1].rstrip(",")) & data:
                        out.append((i + 1, line.strip(), t))
                        break
                waits += 1
            if waits >= 2:
                break
    return out


def test_scanner_finds_the_pattern():
    bad = "\tbuffer_store_dwordx4 v[0:3], v72, s[16:19], s57 offen\n\tv_add_u32_e32 v0, 0x1000, v73\n"
    ok = ("\tbuffer_store_dwordx4 v[0:3], v72, s[16:19], 0 offen\n\tv_mov_b32_e32 v6, v4\n"
          "\tv_mov_b32_e32 v7, v5\n\tv_add_u32_e32 v0, 0x800, v73\n")
    nop = "\tbuffer_store_dwordx4 v[0:3], v72, s[16:19], 0 offen\n\ts_nop 1\n\tv_add_u32_e32 v0, 0x800, v73\n"
    assert len(hazards(bad)) == 1
    assert hazards(ok) == [] and hazards(nop) == []


def test_wide_buffer_stores_use_a_zero_soffset():
    calls = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hpp", ".hip", ".cpp")):
            text = open(os.path.join(CSRC, name)).read()
            calls += [(name, m.start()) for m in re.finditer(r"raw_buffer_store_b(96|128)\(", text)]
    assert calls, "no 16-byte buffer stores found"
    text = open(os.path.join(CSRC, "qs_sweep.hpp")).read()
    body = text[text.index("void buf_store_f4("):]
    body = body[:body.index("\n}\n")]
    assert re.search(r"raw_buffer_store_b128\([^;]*, r, off, 0, 0\)", body)
    assert len(calls) == 1, calls  # the only one is buf_store_f4's


@pytest.fixture(scope="module")
def assembly():
    """name -> gfx950 assembly of the per-size sweep kernels and two compiled plans."""
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not installed")
    tmp = tempfile.mkdtemp(prefix="qs_isa_")
    try:
        jobs = [("sizes%d" % n, [HIPCC] + FLAGS + ["-DQS_N=%d" % n, os.path.join(CSRC, "qs_sizes.hip")])
                for n in (256, 512, 1024, 2048, 4096)]
        jobs += [("mix%d" % n, [HIPCC] + FLAGS + ["-DQS_MIX_N=%d" % n, os.path.join(CSRC, "qs_mix.hip")])
                 for n in (400, 1280)]

        def run(job):
            name, cmd = job
            out = os.path.join(tmp, name + ".s")
            subprocess.run(cmd + ["-o", out], check=True, capture_output=True)
            return name, open(out).read()

        with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as ex:
            return dict(ex.map(run, jobs))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def test_sweep_kernels_have_no_store_data_hazard(assembly):
    found = {name: h for name, h in ((n, hazards(a)) for n, a in assembly.items()) if h}
    assert not found, found


def test_lds_copies_are_range_checked(assembly):
    """The 512^2 row sweep copies state rows into LDS with buffer loads; each carries its
    whole offset in the VGPR offset (soffset 0), which the resource's range check covers,
    so a copy's tail past the last row reads zeros instead of memory past the buffer."""
    copies = [line.strip() for a in assembly.values() for line in a.split("\n")
              if re.match(r"\s*buffer_load_\w+ .* lds$", line)]
    assert any("buffer_load_dwordx4" in c for c in copies), "no LDS copies in the 512^2 row sweep"
    bad = [c for c in copies if not re.match(r"buffer_load_\w+ v\d+, s\[\d+:\d+\], 0 offen lds$", c)]
    assert not bad, bad
