#!/usr/bin/env python3
"""Scan gfx950 kernel assembly for spill writes placed under a branch's
lane mask at a join.

The AMDGPU backend lowers a per-lane branch as: save exec, narrow it to the
lanes that take the branch, skip the block when none do, and restore exec
at the join.  A value merged at the join (assigned on both sides) and then
spilled must be written for every lane; a spill store or AGPR copy emitted
in the join block *before* exec is restored covers only the branch's lanes,
and the other lanes later reload a stale copy.  (An execz test not preceded
by a narrowing of exec is a long-branch relaxation, not a lane mask, and is
not scanned.)  Two builds of these kernels
showed exactly that (every cell of a kernel variant went wrong).  This check
flags any scratch store or AGPR write between an execz-branch target label
and the next write of exec, branch or label.

Usage: check_exec_spills.py file.s [...]; exit status 1 if any is flagged.
"""
import re
import sys

# anything that writes exec, a branch or a label ends the join block
BARRIER = re.compile(r"s_\w+_saveexec_b64|s_\w+_b64\s+exec,|s_cbranch|^\.LBB")


NARROW = re.compile(r"s_(and|andn2|or|xor)_saveexec_b64|s_(mov|and|andn2|xor)_b64\s+exec,")


def scan(path):
    lines = open(path).read().split("\n")
    targets = set()
    for k, ln in enumerate(lines):
        mo = re.search(r"s_cbranch_execz\s+(\S+)", ln)
        if not mo:
            continue
        # a branch around a block run by a subset of lanes: exec was narrowed
        # just before (an execz test without that is a branch relaxation)
        prev = [x for x in lines[max(0, k - 8):k] if x.strip() and not x.strip().startswith(";")]
        if any(NARROW.search(x) for x in prev[-4:]):
            targets.add(mo.group(1))
    found = []
    for i, ln in enumerate(lines):
        if not ln.startswith(".LBB") or ln.split(":")[0].strip() not in targets:
            continue
        for j in range(i + 1, min(len(lines), i + 60)):
            if BARRIER.search(lines[j]):
                break
            if "scratch_store" in lines[j] or "v_accvgpr_write" in lines[j]:
                found.append((j + 1, lines[j].strip()))
    return found


def main(paths):
    bad = 0
    for p in paths:
        found = scan(p)
        bad += len(found)
        print("%s: %d" % (p, len(found)))
        for line, text in found[:8]:
            print("  line %d: %s" % (line, text))
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
