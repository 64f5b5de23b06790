"""Writes plugin_interface_bm3d_v1.json: the BM3D signature and the 900 BM3D
creation probes (data, not code) of the reference's interface golden
(tests/data/plugin_interface_v1.json, schema nssfactory.plugin_interface.v1).

Run in the build container, where the reference tree is mounted:
    python tests/golden/extract_interface.py [/root/reference]
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "tests", "data", "plugin_interface_v1.json")
    with open(src) as f:
        golden = json.load(f)
    cases = {k: v for k, v in golden["cases"].items() if k.split("|")[0] == "BM3D"}
    out = {"schema": golden["schema"], "source": "tests/data/plugin_interface_v1.json (BM3D entries)",
           "signature": golden["functions"]["BM3D"], "cases": cases}
    dst = os.path.join(HERE, "plugin_interface_bm3d_v1.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True, ensure_ascii=False)
        f.write("\n")
    print(f"wrote {dst}: {len(cases)} cases")


if __name__ == "__main__":
    main()
