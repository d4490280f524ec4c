"""Write tests/golden/fortran_thunk_calls.txt: the names of the C functions
that the reference's Fortran thunks (src/fmulti_bspline.c) call, one per
line.  Run in a checkout that has the reference at /root/reference."""
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/fmulti_bspline.c"

names = sorted(set(re.findall(r"\b((?:create|set|eval|destroy)_\w+)\s*\(", open(SRC).read())))
with open(os.path.join(HERE, "fortran_thunk_calls.txt"), "w") as f:
    f.write("\n".join(names) + "\n")
print(len(names), "names")
