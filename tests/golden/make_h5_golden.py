"""Write z0.001t3.16_v2.params.txt: the /params group of the reference's seeded run file
(output/z0.001t3.16_v2.h5, written by initHDF5Output) as the installed `h5dump` prints it.

Run in a container that holds the reference tree; the text it writes is the committed fixture.
"""
import os
import subprocess

REF = "/root/reference"
H5 = os.path.join(REF, "output", "z0.001t3.16_v2.h5")
H5DUMP = "/opt/conda/bin/h5dump"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "z0.001t3.16_v2.params.txt")


def main():
    txt = subprocess.run([H5DUMP, "-g", "/params", H5], capture_output=True, text=True, check=True).stdout
    # the first line names the file
    open(OUT, "w").write(txt.split("\n", 1)[1])


if __name__ == "__main__":
    main()
