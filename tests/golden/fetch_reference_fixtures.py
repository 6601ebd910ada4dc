"""Copy the reference's own sphere-operator golden vectors into the repo.

The files test/unit_tests/inputs/{gradient,divergence,vorticity}_sphere_np4.in
of the reference are plain-text data (inputs, derivative matrix, metric terms
and the Fortran result for one element and one level).  They are copied
verbatim, with a SHA256 manifest, so the oracle can be checked without the
reference mounted.  Run by hand where the reference is available:

    python tests/golden/fetch_reference_fixtures.py [/path/to/reference]
"""
import hashlib
import json
import os
import shutil
import sys

NAMES = ["gradient_sphere_np4.in", "divergence_sphere_np4.in", "vorticity_sphere_np4.in"]


def main(root):
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_inputs")
    os.makedirs(here, exist_ok=True)
    manifest = {}
    for n in NAMES:
        src = os.path.join(root, "test/unit_tests/inputs", n)
        shutil.copyfile(src, os.path.join(here, n))
        manifest[n] = {"source": "test/unit_tests/inputs/" + n,
                       "sha256": hashlib.sha256(open(src, "rb").read()).hexdigest()}
    json.dump(manifest, open(os.path.join(here, "MANIFEST.json"), "w"), indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
