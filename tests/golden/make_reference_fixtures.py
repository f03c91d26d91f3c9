"""Write tests/golden/reference_tests.json from the reference's own test data.

Reads the depth images held by the reference's tests (test/TestImages.cpp, test/MinMaxTest.cpp)
as text, and records the known answers those tests assert (test/*.cpp, cited per entry).
Run in a container where the reference is mounted:  python tests/golden/make_reference_fixtures.py
The JSON it writes is committed; tests never read the reference at run time.
"""
import json
import os
import re

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tests.json")


def _floats(block: str):
    return [float(v) for v in re.findall(r"-?\d+(?:\.\d*)?(?:[eE]-?\d+)?", block)]


def _array(src: str, name: str):
    m = re.search(name + r"\s*\{(.*?)\};", src, re.S)
    return _floats(m.group(1))


def main():
    images = open(os.path.join(REF, "test/TestImages.cpp")).read()
    minmax = open(os.path.join(REF, "test/MinMaxTest.cpp")).read()
    d8, d16, d32 = (_array(images, n) for n in ("depths8x8", "depths16x16", "depths32x32"))
    assert (len(d8), len(d16), len(d32)) == (64, 256, 1024)
    mm8 = _floats(re.search(r"img\.setAll\(vector<float>\{(.*?)\}\);", minmax, re.S).group(1))
    mm4 = _floats(re.search(r"img4x4\.setAll\(std::vector<float>\{(.*?)\}\);", minmax, re.S).group(1))
    assert (len(mm8), len(mm4)) == (64, 16)

    s8, s16, s32 = 2.0 / 8, 2.0 / 16, 2.0 / 32
    fx = {
        "source": "TobiasRp/cpvs test/ (TestImages.cpp, MinMaxTest.cpp, CompressedShadowTest.cpp, CompressedShadowUtilTest.cpp)",
        "images": {"depths8x8": d8, "depths16x16": d16, "depths32x32": d32,
                   "minmax8x8": mm8, "minmax4x4": mm4},
        # MinMaxTest.cpp:41-90 -- [image, level, x, y, which, expected, tolerance]
        "minmax": [
            ["minmax8x8", "numLevels", 4],
            ["minmax8x8", 0, 0, 0, "min", 0.0, 0], ["minmax8x8", 0, 0, 0, "max", 0.0, 0],
            ["minmax8x8", 1, 0, 0, "min", 0.0, 0], ["minmax8x8", 1, 0, 0, "max", 0.9, 0],
            ["minmax8x8", 2, 0, 0, "max", 1.0, 0],
            ["minmax8x8", 3, 0, 0, "max", 1.0, 0], ["minmax8x8", 3, 0, 0, "min", 0.0, 0],
            ["minmax4x4", 1, 0, 0, "min", 0.0, 0], ["minmax4x4", 1, 0, 0, "max", 0.0, 0],
            ["minmax4x4", 1, 1, 0, "max", 1.0, 0],
            ["depths32x32", "numLevels", 6],
            ["depths32x32", 4, 0, 1, "min", 0.673203, 1e-6], ["depths32x32", 4, 0, 0, "min", 0.63008, 1e-6],
            ["depths32x32", 4, 1, 0, "min", 0.63008, 1e-6], ["depths32x32", 4, 1, 1, "min", 0.700469, 1e-6],
            ["depths32x32", 2, 1, 1, "min", 0.63008, 1e-6],
        ],
        # CompressedShadowUtilTest.cpp:14-19
        "childmask": [["depths8x8", 1, [2, 0, 0], 0x88AA]],
        # CompressedShadowUtilTest.cpp:21-27
        "num_children": [[2, 1], [10, 2]],
        # CompressedShadowUtilTest.cpp:29-65 -- [mask, parent, expected first child]
        "child_coordinates": [[2, [0, 0, 0], [0, 0, 0]], [8, [0, 0, 0], [2, 0, 0]],
                              [0x20, [0, 0, 0], [0, 2, 0]], [0x80, [0, 0, 0], [2, 2, 0]],
                              [0x200, [0, 0, 0], [0, 0, 2]], [0x8000, [0, 0, 0], [2, 2, 2]]],
        # CompressedShadowUtilTest.cpp:67-81
        "is_equal_subtree": [[[0xAAAA, 10, 42, 0, 0, 1, 2, 3, 4], [0xAAAA, 10, 42, 0, 0, 1, 2, 3, 4], True],
                             [[0xAAAA, 10, 42, 0, 0, 1, 2, 3, 4], [0xAAA0, 10, 42, 0, 0, 1, 2, 3, 4], False],
                             [[0xAAAA, 10, 42, 0, 0, 1, 2, 3, 4], [0xAAAA, 10, 42, 0, 0, 1, 2, 3, 0], False]],
        # CompressedShadowUtilTest.cpp:83-97 -- level words, expected mapping pairs, expected res[9]
        "merge_level": {"words": [0xAAAA, 10, 42, 0, 0, 1, 2, 3, 4,
                                  0xAAA0, 10, 0, 0, 0, 0, 0, 0, 0,
                                  0xAAAA, 10, 42, 0, 0, 1, 2, 3, 4],
                        "mapping": [[0, 0], [9, 9]], "res9": 0xAAA0},
        # CompressedShadowTest.cpp:50-152 -- [image, ndc, tryLeafmasks, expected (0 SHADOW, 1 VISIBLE)]
        "traverse": [
            ["depths8x8", [-1, 1, -1], False, 0], ["depths8x8", [-1, -1, 0], False, 1],
            ["depths8x8", [1, 1, 0], False, 0], ["depths8x8", [0, 0, 0], False, 0],
            ["depths8x8", [0.5, 0.5, 0.5], False, 0], ["depths8x8", [1, 1, -0.75], False, 1],
            ["depths8x8", [1, 1, 0.5], False, 0], ["depths8x8", [1, -1, 0.99], False, 1],
            ["depths8x8", [0.7, -1, 0.99], False, 0], ["depths8x8", [0.45, -1, -0.9], False, 0],
            ["depths8x8", [0.1, -1.0, -0.9], False, 0],
            ["depths16x16", [1, 1, 0], True, 1], ["depths16x16", [-1, -1, -1.0 + 3 * s16], True, 1],
            ["depths16x16", [-1, -s16, -1.0 + 3 * s16], True, 1], ["depths16x16", [-1, 1, -1], True, 0],
            ["depths16x16", [0, 0, -0.9], True, 1], ["depths16x16", [0 + s16, 0, 0.9], True, 0],
            ["depths16x16", [0 + 3 * s16, 0, -0.9], True, 1],
            ["depths32x32", [1, 1, 0], True, 1],
            ["depths32x32", [-1.0 + 7 * s32, 1.0 - 7 * s32, 0.5], True, 1],
            ["depths32x32", [-1.0 + 7 * s32, 1.0 - 7 * s32, 0.6], True, 0],
            ["depths32x32", [-1.0 + 7 * s32, -1.0 + 8 * s32, 0.3], True, 0],
        ],
        # CompressedShadowTest.cpp:138-151 -- sweeps: (x/32, y/32, z) -> ndc = 2v-1
        "sweeps": [
            {"image": "depths32x32", "x": [0, 31], "y": [0, 31], "z": 0.59, "expected": 1},
            {"image": "depths32x32", "x": [7, 27], "y": [8, 26], "z": 0.79, "expected": 0},
        ],
    }
    del s8
    with open(OUT, "w") as f:
        json.dump(fx, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
