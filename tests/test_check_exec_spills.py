"""The assembly check for spill writes under a branch's lane mask: it flags
the pattern two miscompiled builds of the kernels showed and passes the
ordinary if/else and join shapes."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
spec = importlib.util.spec_from_file_location(
    "check_exec_spills",
    os.path.join(ROOT, "radm2-solver-multi-gpu_amd", "codegen", "check_exec_spills.py"))
C = importlib.util.module_from_spec(spec)
spec.loader.exec_module(C)

# a merged value (3.0 on one side, a load on the other) spilled at the join
# before exec is restored: lanes that skipped the load keep a stale copy
BAD = """\tv_mov_b32_e32 v4, 0
\tv_mov_b32_e32 v5, 0x40080000
\ts_and_saveexec_b64 s[0:1], s[4:5]
\ts_cbranch_execz .LBB0_69
; %bb.68:
\tglobal_load_dwordx2 v[4:5], v[4:5], off offset:376
.LBB0_69:                               ;   in Loop: Header=BB0_6 Depth=1
\ts_waitcnt vmcnt(0)
\tscratch_store_dwordx2 off, v[4:5], off offset:648 ; 8-byte Folded Spill
\ts_or_b64 exec, exec, s[0:1]
"""
BAD_AGPR = BAD.replace("scratch_store_dwordx2 off, v[4:5], off offset:648 ; 8-byte Folded Spill",
                       "v_accvgpr_write_b32 a56, v4")
# the spill after exec is restored, and a store inside an else block
GOOD = """\ts_and_saveexec_b64 s[0:1], vcc
\ts_xor_b64 s[0:1], exec, s[0:1]
\ts_cbranch_execz .LBB0_9
; %bb.8:
\tv_mov_b32_e32 v4, 1
.LBB0_9:
\ts_andn2_saveexec_b64 s[0:1], s[0:1]
\ts_cbranch_execz .LBB0_13
; %bb.10:
\tscratch_store_dwordx2 off, v[10:11], off offset:608 ; 8-byte Folded Spill
.LBB0_13:
\ts_or_b64 exec, exec, s[0:1]
\tscratch_store_dwordx2 off, v[4:5], off offset:648 ; 8-byte Folded Spill
"""


# a long-branch relaxation: an execz test with exec untouched before it
RELAXED = """\tv_fmac_f64_e32 v[34:35], v[26:27], v[26:27]
\ts_cbranch_execz .LBB0_165
\ts_getpc_b64 s[98:99]
\ts_setpc_b64 s[98:99]
.LBB0_165:
\tbuffer_load_dwordx2 v[0:1], v246, s[28:31], s0 offen
\tscratch_store_dwordx2 off, v[0:1], off offset:1180 ; 8-byte Folded Spill
"""


def _scan(tmp_path, text):
    p = tmp_path / "k.s"
    p.write_text(text)
    return C.scan(str(p))


def test_flags_spill_under_branch_mask(tmp_path):
    assert len(_scan(tmp_path, BAD)) == 1
    assert len(_scan(tmp_path, BAD_AGPR)) == 1
    assert C.main([str(tmp_path / "k.s")]) == 1


def test_passes_else_blocks_and_restored_joins(tmp_path):
    assert _scan(tmp_path, RELAXED) == []
    assert _scan(tmp_path, GOOD) == []
    assert C.main([str(tmp_path / "k.s")]) == 0
