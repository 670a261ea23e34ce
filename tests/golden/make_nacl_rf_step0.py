"""Write the NaCl reaction-field step-0 force fixture from the reference's own files.

Source data (reference, read as text): examples/nacl/nacl.cfg (4096 ions, the
input configuration) and examples/nacl/x (the log of a run on that file: a
header line "sim: nve 100 <reacf> <rc^2> ..." and "fld: ... <lj1> <lj2>"
giving the constants, then "FF: i fx fy fz", the step-0 forces of every ion).

The logged build reduced each atom's force correctly only in the first thread
of each thread block (block_size 128 in the log), so the fixture keeps the
atoms i with i % 128 == 0.  Those 32 rows, the input configuration and the
constants are written to nacl_rf_step0/.  Run in a checkout where the
reference sits at /root/reference:  python tests/golden/make_nacl_rf_step0.py
"""
import os
import re
import shutil

REF = "/root/reference/examples/nacl"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nacl_rf_step0")
BLOCK = 128


def main():
    os.makedirs(OUT, exist_ok=True)
    shutil.copyfile(os.path.join(REF, "nacl.cfg"), os.path.join(OUT, "nacl.cfg"))
    sim = fld = None
    rows = {}
    with open(os.path.join(REF, "x")) as f:
        for line in f:
            if line.startswith("sim:") and sim is None:
                sim = line.split()
            elif line.startswith("fld:") and fld is None:
                fld = line.split()
            elif line.startswith("FF:"):
                t = line.split()
                i = int(t[1])
                if i % BLOCK == 0 and i not in rows:     # first block of FF lines = step 0
                    rows[i] = t[2:5]
    reacf, rc2 = float(sim[3]), float(sim[4])
    lj1, lj2 = float(fld[6]), float(fld[7])
    with open(os.path.join(OUT, "forces.txt"), "w") as f:
        f.write("# step-0 forces (kJ/mol/A) of ions i % 128 == 0, examples/nacl/x\n")
        f.write(f"# reacf {reacf} rc2 {rc2} lj1 {lj1} lj2 {lj2}\n")
        for i in sorted(rows):
            f.write(f"{i} {' '.join(rows[i])}\n")
    print("wrote", len(rows), "rows to", OUT)


if __name__ == "__main__":
    main()
