"""Writes the golden fixtures the tests compare against.

Reads the reference's own test data (its .hvm programs and the `run@*.snap`
insta snapshots produced by its test harness, tests/tests.rs:28-42) from
/root/reference and writes:

  tests/golden/programs/<group>/<name>.hvm   the input programs, verbatim
  tests/golden/run_snapshots.json            per program: expected normal form
                                             and the five rewrite counts
  tests/golden/pre_reduce_snapshots.json     per program (`pre_reduce_run@*`,
                                             tests/tests.rs:44-55): the normal
                                             form after pre-reducing every
                                             definition but main, the
                                             pre-reduce counts and the run's
                                             counts; plus the two count sets
                                             of test_fast_pre_reduce
                                             (tests/transform.rs:14-46)

Run once in the build container (the reference tree is not present on the GPU
machine; the fixtures are committed).  The programs are Apache-2.0 data files
of the reference repository (LICENSE there).
"""
import json
import os
import re
import shutil
import sys

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_snapshot(text):
    # insta format: '---' header block '---' then the body
    parts = text.split("---\n")
    body = parts[-1]
    lines = body.rstrip("\n").split("\n")
    counts = {}
    nf_lines = []
    for ln in lines:
        m = re.match(r"^(RWTS|- ANNI|- COMM|- ERAS|- DREF|- OPER)\s*:\s*([0-9_]+)$", ln)
        if m:
            counts[m.group(1).replace("- ", "").lower()] = int(m.group(2).replace("_", ""))
        else:
            nf_lines.append(ln)
    return "\n".join(nf_lines), counts


def parse_pre_reduce_snapshot(text):
    # body: normal form, "pre-reduce:", six count lines, "run:", six count lines
    body = text.split("---\n")[-1].rstrip("\n")
    head, rest = body.split("\npre-reduce:\n")
    pre, run = rest.split("run:\n")
    def counts(block):
        out = {}
        for ln in block.strip("\n").split("\n"):
            m = re.match(r"^(RWTS|- ANNI|- COMM|- ERAS|- DREF|- OPER)\s*:\s*([0-9_]+)$", ln)
            if not m:
                raise ValueError("unexpected count line: " + ln)
            out[m.group(1).replace("- ", "").lower()] = int(m.group(2).replace("_", ""))
        return out
    return head, counts(pre), counts(run)


def fast_pre_reduce_counts():
    # The inline snapshot of test_fast_pre_reduce (tests/transform.rs:31-46):
    # "Fast:" then six count lines, "Slow:" then six count lines.
    src = open(os.path.join(REF, "tests", "transform.rs")).read()
    block = src[src.index("Fast:"):src.index('"###')]
    out, cur = {}, None
    for ln in block.split("\n"):
        ln = ln.strip()
        if ln in ("Fast:", "Slow:"):
            cur = out.setdefault(ln[:-1].lower(), {})
            continue
        m = re.match(r"^(RWTS|- ANNI|- COMM|- ERAS|- DREF|- OPER)\s*:\s*([0-9_]+)$", ln)
        if m and cur is not None:
            cur[m.group(1).replace("- ", "").lower()] = int(m.group(2).replace("_", ""))
    return out


def main():
    if not os.path.isdir(REF):
        print("reference tree not present; fixtures are already committed", file=sys.stderr)
        return 1
    out_prog = os.path.join(HERE, "programs")
    snaps = {}
    snapdir = os.path.join(REF, "tests", "snapshots")
    for fn in sorted(os.listdir(snapdir)):
        if not fn.startswith("run@") or not fn.endswith(".snap"):
            continue
        key = fn[len("run@"):-len(".snap")]  # e.g. tests::programs::stress_tests::boom
        rel = key.replace("::", "/") + ".hvm"
        src = os.path.join(REF, rel)
        if not os.path.exists(src):
            continue
        dst = os.path.join(out_prog, rel)
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        shutil.copyfile(src, dst)
        nf, counts = parse_snapshot(open(os.path.join(snapdir, fn)).read())
        snaps[key] = {"program": os.path.relpath(dst, HERE), "normal_form": nf, "rwts": counts}
    with open(os.path.join(HERE, "run_snapshots.json"), "w") as f:
        json.dump(snaps, f, indent=1, sort_keys=True)
    pre = {}
    for fn in sorted(os.listdir(snapdir)):
        if not fn.startswith("pre_reduce_run@") or not fn.endswith(".snap"):
            continue
        key = fn[len("pre_reduce_run@"):-len(".snap")]
        if key not in snaps:
            continue
        nf, pre_counts, run_counts = parse_pre_reduce_snapshot(open(os.path.join(snapdir, fn)).read())
        pre[key] = {"program": snaps[key]["program"], "normal_form": nf, "pre_rwts": pre_counts,
                    "run_rwts": run_counts}
    fast = fast_pre_reduce_counts()
    pre_doc = {"snapshots": pre,
               "fast_pre_reduce": {"program": "programs/tests/programs/heavy_pre_reduction.hvm",
                                   "removed": "black_box",
                                   "fast": {"reduce": ["expensive", "main_fast"], "entry": "main_fast",
                                            "rwts": fast["fast"]},
                                   "slow": {"reduce": ["expensive_1", "expensive_2", "main_slow"],
                                            "entry": "main_slow", "rwts": fast["slow"]}}}
    with open(os.path.join(HERE, "pre_reduce_snapshots.json"), "w") as f:
        json.dump(pre_doc, f, indent=1, sort_keys=True)
    # tests/programs/list_put_got.hvm is also read by the count-vector tests
    # (tests/lists.rs); it is copied above with its snapshot.
    print(f"wrote {len(snaps)} run and {len(pre)} pre-reduce snapshot fixtures")
    return 0


if __name__ == "__main__":
    sys.exit(main())
