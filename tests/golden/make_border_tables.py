"""Extract the border-copy tables of VulkanExp's probesCopyBorders.comp as a JSON fixture.

Reads /root/reference/src/shaders/probesCopyBorders.comp (lines 21-220: irradianceCopiesDst/Src
with 28 entries, depthCopiesDst/Src with 60 entries) and writes border_tables.json beside this
script: {"irradiance": {"dst": [[x, y], ...], "src": [...]}, "depth": {...}}.  Only the
coordinate pairs are kept.
"""
import json
import os
import re

SRC = "/root/reference/src/shaders/probesCopyBorders.comp"


def table(text: str, name: str):
    m = re.search(name + r"\[[^\]]*\]\s*=\s*\{(.*?)\};", text, re.S)
    return [[int(a), int(b)] for a, b in re.findall(r"ivec2\(\s*(-?\d+)\s*,\s*(-?\d+)\s*\)", m.group(1))]


def main():
    text = open(SRC).read()
    out = {
        "source": "src/shaders/probesCopyBorders.comp:21-220",
        "irradiance": {"res": 8, "dst": table(text, "irradianceCopiesDst"), "src": table(text, "irradianceCopiesSrc")},
        "depth": {"res": 16, "dst": table(text, "depthCopiesDst"), "src": table(text, "depthCopiesSrc")},
    }
    assert len(out["irradiance"]["dst"]) == len(out["irradiance"]["src"]) == 28
    assert len(out["depth"]["dst"]) == len(out["depth"]["src"]) == 60
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "border_tables.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=None, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    main()
