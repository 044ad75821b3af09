"""Extract the reference's golden outputs into a JSON fixture.

Reads the STDOUT templates shipped with the reference
(src/example/template_{cpu,gpu,Eigen_cpu}/STDOUT*) and writes
tests/golden/templates.json: residual histories (one list per printed
column), the max-error lines and the eigenvalue lines.  The fixture is data
only (numbers parsed from the reference's outputs); run this script in the
build container where the reference tree is mounted:

    python tests/golden/extract_templates.py /root/reference
"""
import json
import os
import re
import sys

FILES = [
    "template_cpu/STDOUT_run",
    "template_cpu/STDOUT_run_dd",
    "template_cpu/STDOUT_run_ddm",
    "template_cpu/STDOUT_run_ddm_dd",
    "template_gpu/STDOUT_run",
    "template_gpu/STDOUT_run_ddm",
    "template_Eigen_cpu/STDOUT",
    "template_Eigen_cpu/STDOUT_ddm",
]

NUM = r"[-+]?\d*\.\d+(?:[Ee][-+]?\d+)?|[-+]?\d+(?:[Ee][-+]?\d+)?"


def parse(path):
    blocks = [[]]
    extras = {}
    with open(path) as f:
        for line in f:
            s = line.strip()
            if s.startswith("-----"):
                blocks.append([])
                continue
            m = re.match(r"^(?:PARCEL : @res@\s+)?(\d+)\s+((?:\s*(?:%s))+)$" % NUM, s)
            if m and ("@res@" in s or re.match(r"^\d+\s", s)):
                vals = [float(v) for v in re.findall(NUM, m.group(2))]
                blocks[-1].append([int(m.group(1))] + vals)
                continue
            m = re.match(r"^max\(\|numerical-exact\|\)\s+=\s+(%s)" % NUM, s)
            if m:
                extras.setdefault("max_abs_err", []).append(float(m.group(1)))
            m = re.match(r"^max\(\|numerical-exact\|\) / max\(\|exact\|\)\s+=\s+(%s)" % NUM, s)
            if m:
                extras.setdefault("max_rel_err", []).append(float(m.group(1)))
            m = re.match(r"^(Max|Min) Eigen Value\s+=\s+(%s)\s+(%s)" % (NUM, NUM), s)
            if m:
                extras["eig_" + m.group(1).lower()] = [float(m.group(2)), float(m.group(3))]
            m = re.match(r"^Error \|\|Ax-kx\|\|/\|\|kx\|\|\s+=\s+(%s)" % NUM, s)
            if m:
                extras.setdefault("eig_err", []).append(float(m.group(1)))
            m = re.match(r"^(?:solve\s+time|solve time)\s+(%s)" % NUM, s)
            if m:
                extras.setdefault("solve_time_s", []).append(float(m.group(1)))
    blocks = [b for b in blocks if b]
    return {"histories": blocks, "extras": extras}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    base = os.path.join(ref, "src", "example")
    out = {}
    for rel in FILES:
        out[rel] = parse(os.path.join(base, rel))
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "templates.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0)
    for k, v in out.items():
        print(k, [len(b) for b in v["histories"]], sorted(v["extras"]))


if __name__ == "__main__":
    main()
