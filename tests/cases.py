"""The reference's golden render cases (tests/golden/ao_path_golden.json) as inputs."""
import hashlib
import json
import os

GOLDEN = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ao_path_golden.json")))
CASES = GOLDEN["identity_cases"]
TREES = GOLDEN["ploc_trees"]
KNOWN = GOLDEN["reconcile_known_answers_v0_3_0"]
FRAMES = GOLDEN["frame_cases"]
# the v0.3.0 known answers whose view the identity matrix still renders
KNOWN_CASE = {"default": "default-512", "high": "high-512", "low": "low-512", "wide": "wide-512"}


def params(name):
    """CLI defaults (raw/tools/cli_params.hpp:14-43) overridden by the case's arguments."""
    p = dict(w=256, h=256, eye=(4.0, 4.0, 6.0), target=(0.0, 1.0, 0.0), fovy=0.9, threads=1, prev_eye=None, rt=True)
    a = (CASES.get(name) or FRAMES[name])["args"]
    i = 0
    while i < len(a):
        k = a[i]
        if k == "--width":
            p["w"] = int(a[i + 1])
        elif k == "--height":
            p["h"] = int(a[i + 1])
        elif k == "--eye":
            p["eye"] = tuple(float(x) for x in a[i + 1].split(","))
        elif k == "--target":
            p["target"] = tuple(float(x) for x in a[i + 1].split(","))
        elif k == "--fovy":
            p["fovy"] = float(a[i + 1])
        elif k == "--threads":
            p["threads"] = int(a[i + 1])
        elif k == "--prev-eye":
            p["prev_eye"] = tuple(float(x) for x in a[i + 1].split(","))
        elif k == "--no-rt":
            p["rt"] = False
        i += 2 if k.startswith("--") and k not in ("--no-rt", "--gpu") else 1
    return p


def sha(b):
    return hashlib.sha256(b).hexdigest()
