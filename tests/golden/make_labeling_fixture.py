"""Writes tests/golden/labeling_kat.json from the reference's labelling tests.

Source: src/test/sourcefind/sourcefind_test_Labeling.cpp:73-675 of the
reference repository.  Each of the twelve TEST_F blocks holds an 8x8 input
image (symbols S = 10, I = -10, n = 0), the connectivity, the expected label
map for the non-FFTSHIFT build (the production build) and the expected island
count.  Only that data is extracted; the run parameters are those of the test
fixture's SetUp() (:34-46): detection 4, analysis 3, rms_est 1, ZEROMEDIAN,
find_negative_sources, generate_labelmap false, source_min_area 1.

Usage: python tests/golden/make_labeling_fixture.py [reference_root]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(REF, "src/test/sourcefind/sourcefind_test_Labeling.cpp")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "labeling_kat.json")

SYM = {"S": 10.0, "I": -10.0, "n": 0.0}


def parse_matrix(text, numeric):
    rows = re.findall(r"\{([^{}]*)\}", text)
    out = []
    for r in rows:
        toks = [t.strip() for t in r.split(",") if t.strip()]
        out.append([int(t) if numeric else SYM[t] for t in toks])
    return out


def main():
    src = open(SRC).read()
    blocks = re.split(r"TEST_F\(SourceFindLabeling,\s*", src)[1:]
    cases = []
    for b in blocks:
        name = b[: b.index(")")]
        conn4 = re.search(r"ccl_4connectivity\s*=\s*(true|false)", b).group(1) == "true"
        img_txt = re.search(r"img\s*=\s*\{(.*?)\};", b, re.S).group(1)
        nofft = b[b.index("#ifndef FFTSHIFT"): b.index("#else")]
        exp_txt = re.search(r"expected_map\s*=\s*\{(.*?)\};", nofft, re.S).group(1)
        total = int(re.search(r"expected_total_islands\s*=\s*(\d+)", b).group(1))
        cases.append({
            "name": name,
            "ccl_4connectivity": conn4,
            "image": parse_matrix(img_txt, False),
            "expected_map_nonfftshift": parse_matrix(exp_txt, True),
            "expected_total_islands": total,
        })
    doc = {
        "source": "src/test/sourcefind/sourcefind_test_Labeling.cpp:73-675 (expected maps of the #ifndef FFTSHIFT branches)",
        "params": {"detection_n_sigma": 4.0, "analysis_n_sigma": 3.0, "rms_est": 1.0, "median_method": "ZEROMEDIAN",
                   "find_negative_sources": True, "generate_labelmap": False, "source_min_area": 1,
                   "sigma_clip_iters": 5},
        "procedure": "the test fftshifts the image before source finding and fftshifts the label map back "
                     "(:48-69); then repeats with the image negated and expects the negated map",
        "cases": cases,
    }
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
