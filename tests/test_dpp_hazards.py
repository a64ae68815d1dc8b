"""The broadcast-row flips are inline-asm v_fmac_f64 with DPP row_newbcast, which the
compiler's hazard recognizer does not see. On gfx950 a DPP read of a VGPR needs 2 wait
states after a VALU write of it. This test compiles the kernels that use them to device
assembly (CPU only) and checks that no VALU write of a DPP source register sits in the 2
instructions before the DPP read, and that no DPP read follows a branch target within 2
instructions (where the predecessor is not visible)."""
import concurrent.futures as cf
import os
import re
import subprocess
import tempfile

import pytest

from theboss_amd import build as B

TUS = ("sweep", "glynn_f64", "grad", "sampler")


def _vgprs(tok):
    m = re.match(r"v\[(\d+):(\d+)\]$", tok)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"v(\d+)$", tok)
    return {int(m.group(1))} if m else set()


def scan(asm: str):
    """(number of DPP broadcast reads, list of hazards) of one assembly listing."""
    total, bad = 0, []
    prev, since_label, func = [], 99, "?"
    for line in asm.split("\n"):
        s = line.strip()
        if not s or s.startswith((".", ";")):
            continue
        m = re.match(r"^([\w.$]+):", s)
        if m:
            if m.group(1).startswith("_Z"):
                func = m.group(1)
            prev, since_label = [], 0
            continue
        op = s.split()[0]
        if "_dpp" in op and "row_newbcast" in s:
            total += 1
            if since_label < 2:
                bad.append((func, "after branch target", s))
            src0 = _vgprs(s[len(op):].split(",")[1].strip().lstrip("-"))
            for w in prev[-2:]:
                wop = w.split()[0]
                if wop.startswith("v_") and not wop.startswith(("v_readlane", "v_readfirstlane", "v_cmp")):
                    if _vgprs(w[len(wop):].split(",")[0].strip()) & src0:
                        bad.append((func, w, s))
        n = int(s.split()[1], 0) + 1 if op == "s_nop" else 1
        prev.extend([op] * (n - 1) + [s] if op == "s_nop" else [s])
        since_label += n
        prev = prev[-4:]
    return total, bad


def test_scanner_flags_a_planted_hazard():
    asm = "\n".join(["f:", "v_mov_b32 v0, 1", "v_add_f64 v[8:9], v[2:3], v[4:5]",
                     "v_fmac_f64_dpp v[6:7], v[8:9], v[10:11] row_newbcast:1 row_mask:0xf bank_mask:0xf",
                     "s_nop 1", "v_fmac_f64_dpp v[6:7], v[2:3], v[10:11] row_newbcast:2 row_mask:0xf bank_mask:0xf"])
    total, bad = scan(asm)
    assert total == 2 and len(bad) == 1 and "v[8:9]" in bad[0][2]
    total, bad = scan("\n".join(["g:", "v_add_f64 v[8:9], v[2:3], v[4:5]", "s_nop 1",
                                 "v_fmac_f64_dpp v[6:7], v[8:9], v[10:11] row_newbcast:1"]))
    assert total == 1 and not bad


@pytest.mark.timeout(900)
def test_kernels_have_no_dpp_read_after_valu_write():
    try:
        hipcc = B._hipcc()
    except RuntimeError:
        pytest.skip("hipcc not available")
    with tempfile.TemporaryDirectory() as tmp:
        def compile_one(tu):
            out = os.path.join(tmp, tu + ".s")
            cmd = [hipcc, *B.HIPCC_FLAGS, "--cuda-device-only", "-S", "-I", B.CSRC,
                   os.path.join(B.CSRC, tu + ".hip"), "-o", out]
            subprocess.run(cmd, check=True, capture_output=True)
            with open(out) as f:
                return tu, f.read()
        with cf.ThreadPoolExecutor(max_workers=len(TUS)) as ex:
            listings = dict(ex.map(compile_one, TUS))
    for tu, asm in listings.items():
        total, bad = scan(asm)
        assert total > 0, tu  # the broadcast flips are there
        assert not bad, (tu, bad[:3])
