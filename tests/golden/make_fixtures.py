"""Copies the reference's own data fixtures for this path into tests/golden/.

Run in the build container (where the reference is mounted); the copies are
committed so that nothing at test time reads the reference.  They are data
(tables of <m>(t)), not code:
  stdsolution4.txt   tests/stdsolution4.txt            (SP4 <m>(t), t in ns)
  sp5_m.txt          certification/stdproblem5/m.txt   (SP5 Zhang-Li <m>(t), t in s)
"""
import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = {
    "stdsolution4.txt": "tests/stdsolution4.txt",
    "sp5_m.txt": "certification/stdproblem5/m.txt",
}

if __name__ == "__main__":
    for dst, src in FILES.items():
        shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, dst))
        print("copied", src, "->", dst)
