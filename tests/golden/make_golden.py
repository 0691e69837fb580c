"""Extract the small data fixtures this test suite uses from the reference
tree (data files only: parameter files, the pair-index file, and the header
lines / first rows / shape of each committed result field).  Run in a
container where the reference is mounted; writes next to this script."""
import json
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

shutil.copyfile(os.path.join(REF, "xcorr2d_V1.0", "Parameter.dat"), os.path.join(HERE, "Parameter.dat"))
shutil.copyfile(os.path.join(REF, "x64", "Debug", "Parameter.dat"), os.path.join(HERE, "Parameter41.dat"))
shutil.copyfile(os.path.join(REF, "data", "data_organized", "pair index.txt"), os.path.join(HERE, "pair_index.txt"))

res = os.path.join(REF, "data", "data_organized", "result")
out = {}
for name in sorted(os.listdir(res)):
    path = os.path.join(res, name)
    if not name.endswith(".dat"):
        continue
    with open(path, "rb") as f:
        lines = f.read().split(b"\n")
    if lines and lines[-1] == b"":
        lines = lines[:-1]
    out[name] = dict(header=[l.decode() + "\n" for l in lines[:3]], first_rows=[l.decode() + "\n" for l in lines[3:8]],
                     last_row=lines[-1].decode() + "\n", n_lines=len(lines))
with open(os.path.join(HERE, "result_fixtures.json"), "w") as f:
    json.dump(out, f, indent=1)
print("wrote", len(out), "result fixtures")
