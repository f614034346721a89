"""Copy the reference's own regression fixtures for the surface slice loop into
tests/golden/ref/ (data only: decks, bulkP.b and the recorded FP64 rocking curves).

Source: test/test{23,47,521}/ of the reference repository, produced by its authors'
builds (see test/test23/setup.sh and perf-test.sh there), and the one-beam sample
sim-trhepd-rheed/sample/Ge_001_c4x2/ (deck, bulkP.b, the recorded curve, E12.4, and its angular
convolution), and str-orig/sample/T4Al_on_Si_111/output (a recorded curve and its convolution by
that program's tool).  Run once in a container
where the reference is mounted; the GPU hosts only read the copies.
"""
import os
import shutil
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref")
PICK = {
    "test23": ["0.68", "0.1", "0.01"],
    "test47": ["0.68", "0.1"],
    # N = 521: the coarse pair plus the paper's headline steps (sp4 0.068, sp6 0.15, rk4 0.015,
    # test/plots-plus/table.tex:17-19) and the fine steps up to the bench's slice count and beyond
    "test521": ["0.68", "0.47", "0.15", "0.1", "0.068", "0.047", "0.022", "0.015", "0.01"],
}
for case, dzs in PICK.items():
    os.makedirs(os.path.join(DST, case), exist_ok=True)
    for f in ("bulk.txt", "surf.txt", "bulkP.b"):
        shutil.copy(os.path.join(SRC, case, f), os.path.join(DST, case, f))
    shutil.copy(os.path.join(SRC, case, "surf-bulkP-str-accurate.s"), os.path.join(DST, case, "sp4-0.001.s"))
    for m in (4, 5, 6, 7):
        for dz in dzs:
            shutil.copy(os.path.join(SRC, case, "test-results", "perf-%d-%s.txt" % (m, dz)),
                        os.path.join(DST, case, "perf-%d-%s.txt" % (m, dz)))
GE = os.path.join(os.path.dirname(SRC.rstrip("/")), "sim-trhepd-rheed", "sample", "Ge_001_c4x2")
os.makedirs(os.path.join(DST, "ge001"), exist_ok=True)
for f in ("bulk.txt", "surf.txt"):
    shutil.copy(os.path.join(GE, f), os.path.join(DST, "ge001", f))
for f in ("bulkP.b", "surf-bulkP.s", "convolution.txt"):
    shutil.copy(os.path.join(GE, "output", f), os.path.join(DST, "ge001", f))
# the angular convolution of the original program's tool (str-orig/tool/make_convolution.py)
# on its T4 Al / Si(111) sample: the recorded curve and its convolution
T4 = os.path.join(os.path.dirname(SRC.rstrip("/")), "str-orig", "sample", "T4Al_on_Si_111", "output")
os.makedirs(os.path.join(DST, "t4al"), exist_ok=True)
for f in ("surf-bulkP.s", "convolution.txt"):
    shutil.copy(os.path.join(T4, f), os.path.join(DST, "t4al", f))
print("copied to", DST)
