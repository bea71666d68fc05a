#!/usr/bin/env python3
"""Extracts the reference's published golden values for the AO path into
tests/golden/ao_path_golden.json (data only: arguments, hashes and numbers).

Sources (read as JSON text from the reference tree):
  evidence/identity-golden.json             SHA-256 of every CPU render's files (lighting model 2)
  evidence/identity-golden-v1.json          the same under lighting model 1 (frame files only)
  evidence/rt-r2-bvh-gpu-d3d12-rtx4090.json PLOC tree hashes (FNV-1a) and rounds; the R2 check's
                                            rays, hits (100,000 rays a scene) and its controls
  evidence/rt-r2-bvh-cpu.json               PLOC rounds, depth, node counts, SAH
  evidence/wasm-vs-native-512.json          reconcile exact values (pixels, rmse, maxError)
Usage: make_golden.py [/root/reference]
"""
import json
import os
import sys

AO_FILES = ("ao_rt.pfm", "ao_ss.pfm", "mask.pgm", "ao_error.pgm", "ao_rt.pgm", "ao_ss.pgm", "certificate.json")
FRAME_FILES = ("frame.ppm", "frame_hdr.pfm", "channels.json", "receipt.json")


def main(ref):
    ev = lambda p: json.load(open(os.path.join(ref, "evidence", p)))
    ident = ev("identity-golden.json")["cases"]
    ident_v1 = ev("identity-golden-v1.json")["cases"]
    cases, frames = {}, {}
    for name, c in ident.items():
        if "--gpu" in c["args"]:
            continue
        files = c["files"]
        frames[name] = {"args": c["args"], "files": {k: v for k, v in files.items() if k in AO_FILES + FRAME_FILES},
                        "files_lighting_v1": {k: v for k, v in ident_v1[name]["files"].items()
                                              if k in ("frame.ppm", "frame_hdr.pfm")}}
        if "--no-rt" in c["args"]:
            continue
        cases[name] = {"args": c["args"], "files": {k: v for k, v in files.items() if k in AO_FILES}}
    gpu = ev("rt-r2-bvh-gpu-d3d12-rtx4090.json")
    cpu = {r["scene"]: r for r in ev("rt-r2-bvh-cpu.json")["results"]}
    trees = {}
    for s in gpu["scenes"]:
        c = cpu[s["scene"]]
        trees[s["scene"]] = {"triangles": s["triangles"], "tree_hash": s["cpu_tree_hash"], "rounds": s["rounds"],
                             "depth": c["depth"], "sah": c["sah"], "m3_binned_sah": c["m3_binned_sah"],
                             "binned_nodes": c["bvh_node_count"], "hits_100k": c["hits"],
                             "r2_rays": s["rays"], "r2_hits": s["hits"], "r2_sah_ratio": s["sah_ratio"]}
    wasm = ev("wasm-vs-native-512.json")
    known = {k: {f: v["a"][f] for f in ("pixels", "rmse", "maxError", "verdict")} for k, v in wasm.items()
             if isinstance(v, dict) and "a" in v}
    out = {"source": "raw-native v0.8.0 evidence/ (identity-golden, rt-r2-bvh-*, wasm-vs-native-512)",
           "identity_cases": cases, "frame_cases": frames, "ploc_trees": trees, "reconcile_known_answers_v0_3_0": known,
           "control_radius_8_nodes_differing_dense_block": gpu["control_radius_8_nodes_differing"],
           "control_short_far_hits_missed_retro_room": gpu["control_short_far_hits_missed"],
           "default_256_reconcile": {"pixels": 37996, "rmse": 0.129436031}}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ao_path_golden.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
