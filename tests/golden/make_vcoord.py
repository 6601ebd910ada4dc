"""Regenerate the hybrid vertical-coordinate tables shipped with the package.

The numbers are data read from the reference's ASCII coordinate files
(test/vcoord/{acme-72i,acme-72m,cami-26,camm-26}.ascii): each file holds a
count line followed by that many coefficients (interfaces: hyai then hybi;
midpoints: hyam then hybm).  This script is run by hand in a container that
has the reference mounted; the JSON it writes into hommexx_amd/data/ is what
the package loads.

    python tests/golden/make_vcoord.py [/path/to/reference]
"""
import hashlib
import json
import os
import sys

SOURCES = {
    72: ("test/vcoord/acme-72i.ascii", "test/vcoord/acme-72m.ascii"),
    26: ("test/vcoord/cami-26.ascii", "test/vcoord/camm-26.ascii"),
}


def _read_two_blocks(path):
    toks = open(path).read().split()
    out = []
    i = 0
    while i < len(toks) and len(out) < 2:
        n = int(toks[i])
        i += 1
        # skip the comment marker and label ("!", "hyai")
        while not _is_float(toks[i]):
            i += 1
        vals = [float(t) for t in toks[i:i + n]]
        i += n
        out.append(vals)
    return out


def _is_float(t):
    try:
        float(t)
        return True
    except ValueError:
        return False


def main(ref_root):
    out_dir = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                        "hommexx_amd", "data")
    for nlev, (fi, fm) in SOURCES.items():
        pi = os.path.join(ref_root, fi)
        pm = os.path.join(ref_root, fm)
        hyai, hybi = _read_two_blocks(pi)
        hyam, hybm = _read_two_blocks(pm)
        assert len(hyai) == nlev + 1 and len(hybi) == nlev + 1
        assert len(hyam) == nlev and len(hybm) == nlev
        rec = {
            "nlev": nlev,
            "hyai": hyai, "hybi": hybi, "hyam": hyam, "hybm": hybm,
            "source": [fi, fm],
            "sha256": [hashlib.sha256(open(p, "rb").read()).hexdigest() for p in (pi, pm)],
        }
        with open(os.path.join(out_dir, "vcoord_L%d.json" % nlev), "w") as f:
            json.dump(rec, f, indent=1)
        print("wrote vcoord_L%d.json" % nlev)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
