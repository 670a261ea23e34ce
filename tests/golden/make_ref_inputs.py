"""Copy the reference's example input files used as parser fixtures.

They are data files from the reference's examples/ tree (positional .inp,
.fld and .cfg formats, src/simulation.cu:40-316); the tests parse them with
the C++ SimulationInit and the Python reader and compare the two.  Run in a
checkout where the reference sits at /root/reference:
    python tests/golden/make_ref_inputs.py
"""
import os
import shutil

REF = "/root/reference/examples"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_inputs")
FILES = {
    "h2o_wolf_h2o.inp": "h2o/wolf/h2o.inp",
    "h2o_wolf_h2o.fld": "h2o/wolf/h2o.fld",
    "h2o_wolf_h2o.cfg": "h2o/wolf/h2o.cfg",
    "ar_ar.inp": "ar/ar.inp",
    "ar_ar.fld": "ar/ar.fld",
    "nacl_nacl.inp": "nacl/nacl.inp",
    "nacl_nacl.fld": "nacl/nacl.fld",
    "wat_wat.inp": "wat/wat.inp",
    "co2_co2.inp": "co2/co2.inp",
}


def main():
    os.makedirs(OUT, exist_ok=True)
    for dst, src in FILES.items():
        shutil.copyfile(os.path.join(REF, src), os.path.join(OUT, dst))
        os.chmod(os.path.join(OUT, dst), 0o644)
    print("copied", len(FILES), "files to", OUT)


if __name__ == "__main__":
    main()
