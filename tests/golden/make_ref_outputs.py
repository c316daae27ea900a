"""Writes tests/golden/ref_outputs.json from the reference's own CPU
programs, built by `make -C oracle ref` from /root/reference into
oracle/_ref (never copied).  Run in a container that has /root/reference;
the JSON it writes is the committed fixture.

Recorded per program: sha256 and line count of its stdout (the system-test
oracles print one "%f" line per grid element), the scalar results of the
reduction tests, the %f dump of the C++ diffusion baseline after a few
sweeps, and the original Himeno benchmark's p (float bits) and gosa after a
few sweeps on seeded inputs.
"""
import hashlib
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")

TEST_CASES = "/root/reference/tests/system_tests/test_cases"
# Every hand-written REFERENCE-target oracle of the system tests.
SYSTEM_TESTS = sorted(f[:-len(".manual.ref.c")] for f in os.listdir(TEST_CASES)
                      if f.endswith(".manual.ref.c"))
# (size, count) of the C++ diffusion baseline runs.
DIFFUSION_RUNS = [(32, 4), (24, 10)]
# (nx, ny, nz, sweeps, seed) of the original Himeno runs.
HIMENO_RUNS = [(12, 10, 9, 4, 1), (33, 17, 20, 6, 7)]


def run(cmd, cwd=None):
    return subprocess.run(cmd, check=True, stdout=subprocess.PIPE,
                          cwd=cwd).stdout


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"],
                   check=True)
    out = {"system_tests": {}, "diffusion_baseline": {}, "himeno_original": {}}
    for t in SYSTEM_TESTS:
        o = run([os.path.join(REF, t)])
        lines = o.decode().splitlines()
        rec = {"sha256": hashlib.sha256(o).hexdigest(), "lines": len(lines)}
        if len(lines) <= 2:
            rec["values"] = lines
        out["system_tests"][t] = rec
    for size, count in DIFFUSION_RUNS:
        with tempfile.TemporaryDirectory() as d:
            log = run([os.path.join(REF, "diffusion_baseline"), "--count",
                       str(count), "--size", str(size), "--dump"], cwd=d)
            dump = open(os.path.join(d, "diffusion3d_result.baseline.out"),
                        "rb").read()
        acc = [l for l in log.decode().splitlines() if l.startswith("Accuracy")]
        out["diffusion_baseline"][f"{size}x{count}"] = {
            "sha256": hashlib.sha256(dump).hexdigest(),
            "accuracy": acc[0].split(":")[1].strip()}
    for nx, ny, nz, nn, seed in HIMENO_RUNS:
        o = run([os.path.join(REF, "himeno_ref"), str(nx), str(ny), str(nz),
                 str(nn), str(seed)]).decode().splitlines()
        body = "\n".join(o[1:]) + "\n"
        out["himeno_original"][f"{nx}x{ny}x{nz}x{nn}s{seed}"] = {
            "gosa": o[0].split()[1],
            "p_bits_sha256": hashlib.sha256(body.encode()).hexdigest()}
    with open(os.path.join(HERE, "ref_outputs.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps(out, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
