"""Writes the fixtures taken from the reference's shipped inputs (run in the build
container, where the reference tree is present; the outputs are committed as data):

* tube_sphere.dat            -- input/geometry/tube_sphere.dat, byte for byte (60 x 60 x 80
                                voxels, 4-byte dimension header);
* reference_control.json     -- the values of input/simulation_control.txt as parsed by
                                mflbm.config.read_control (every key the reference reads).

    python tests/golden/make_reference_inputs.py [/root/reference]
"""
import json
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "mf-lbm-cuda_amd"))

from mflbm import config  # noqa: E402


def main(ref):
    shutil.copyfile(os.path.join(ref, "input", "geometry", "tube_sphere.dat"), os.path.join(HERE, "tube_sphere.dat"))
    ctrl = config.read_control(os.path.join(ref, "input", "simulation_control.txt"))
    with open(os.path.join(HERE, "reference_control.json"), "w") as fh:
        json.dump(ctrl, fh, indent=1, sort_keys=True)
        fh.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
