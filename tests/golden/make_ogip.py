"""Copies the NICER effective-area file the reference ships
(examples/examples_modeling_tutorial/model_data/nicer_20170601v005.arf) into
tests/golden as a fixture for the OGIP reader.  Run once, in a tree where
the reference is mounted."""
import os
import shutil

SRC = "/root/reference/examples/examples_modeling_tutorial/model_data/nicer_20170601v005.arf"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nicer_20170601v005.arf")

if __name__ == "__main__":
    shutil.copyfile(SRC, DST)
    print(DST, os.path.getsize(DST))
