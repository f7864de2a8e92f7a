"""Extract the golden vectors held by elsa's own tests into tests/golden/*.json.

Run once in a container that has the reference tree (REF defaults to /root/reference):
    python tests/golden/extract_reference_vectors.py
It copies numbers only (expected outputs of the reference tests), never source text:
  luts.json    elsa/projectors/tests/test_Luts.cpp  ProjectedBlobLut (a=2, alpha=10.83, m=2,
               101 entries) and ProjectedBSplineLut (order 3, dim 2, 101 entries)
  bessel.json  elsa/core/tests/test_Bessel.cpp  I_0..I_4 on LinSpaced(size, 0, 4)
  derivatives.json
               elsa/projectors/tests/test_BlobDerivative.cpp  projected blob derivative, a = 1,
               alpha in {2.5, 5, 7.5, 10.83}, m in {1, 2}, 101 values at s = i / 100 (the
               arrays of test_BlobGradientHelper.cpp are checked to be the same numbers);
               elsa/projectors/tests/test_BSplines.cpp  ProjectedBSpline derivative LUT,
               order 3, dim 2, 101 values at x = i / 50
  blob_weights.json
               elsa/projectors/tests/test_BlobVoxelProjector.cpp "Test weights": the projected
               blob (a = 2, alpha = 10.83, m = 2, closed form) at distance i / 25, 51 values
"""
import json
import os
import re
import sys

REF = os.environ.get("REF", "/root/reference")
OUT = os.path.dirname(os.path.abspath(__file__))
NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?"


def numbers(block: str):
    return [float(x) for x in re.findall(NUM, block)]


def brace_block(text: str, start: int) -> str:
    i = text.index("{", start)
    depth = 0
    for j in range(i, len(text)):
        if text[j] == "{":
            depth += 1
        elif text[j] == "}":
            depth -= 1
            if depth == 0:
                return text[i + 1:j]
    raise ValueError("unbalanced braces")


def main():
    luts_src = open(os.path.join(REF, "elsa/projectors/tests/test_Luts.cpp")).read()
    blob_at = luts_src.index('TEST_CASE_TEMPLATE("ProjectedBlobLut: "')
    blob = numbers(brace_block(luts_src, luts_src.index("expected", blob_at)))
    bsp_at = luts_src.index('TEST_CASE_TEMPLATE("ProjectedBSplineLut: "')
    bsp = numbers(brace_block(luts_src, luts_src.index("expected", bsp_at)))
    assert len(blob) == 101 and len(bsp) == 101, (len(blob), len(bsp))
    json.dump({
        "source": "elsa/projectors/tests/test_Luts.cpp",
        "projected_blob": {"radius": 2.0, "alpha": 10.83, "order": 2, "N": 101, "distance": "i/50",
                           "checked": "i < 99", "epsilon": 1e-6, "expected": blob},
        "projected_bspline": {"dim": 2, "order": 3, "N": 101, "distance": "i/50", "checked": "i < 101",
                              "epsilon": 1e-6, "expected": bsp},
    }, open(os.path.join(OUT, "luts.json"), "w"), indent=1)

    bes_src = open(os.path.join(REF, "elsa/core/tests/test_Bessel.cpp")).read()
    tables = {}
    for m in range(5):
        at = bes_src.index(f"std::vector<T> get_bessi{m}()\n{{")
        tables[f"I{m}"] = numbers(brace_block(bes_src, bes_src.index("return", at)))
    json.dump({"source": "elsa/core/tests/test_Bessel.cpp", "x": "LinSpaced(size, 0, 4)", **tables},
              open(os.path.join(OUT, "bessel.json"), "w"), indent=1)
    def blob_tables(path):
        src = open(os.path.join(REF, path)).read()
        body = src[src.index("getExpected(data_t a, data_t alpha, data_t m)\n{"):]
        out, alpha, m, pos = {}, None, None, 0
        pat = re.compile(r"alpha == (?:Approx\()?(" + NUM + r")|m == (\d)|return \{")
        while True:
            hit = pat.search(body, pos)
            if not hit:
                break
            if hit.group(1):
                alpha = float(hit.group(1))
            elif hit.group(2):
                m = int(hit.group(2))
            else:
                block = brace_block(body, hit.start())
                vals = numbers(block)
                if len(vals) == 101 and alpha is not None and m is not None and (alpha, m) not in out:
                    out[(alpha, m)] = vals
            pos = hit.end()
        return out

    deriv = blob_tables("elsa/projectors/tests/test_BlobDerivative.cpp")
    grad = blob_tables("elsa/projectors/tests/test_BlobGradientHelper.cpp")
    assert len(deriv) == 8, sorted(deriv)
    same = sorted(deriv) == sorted(grad) and all(deriv[k] == grad[k] for k in deriv)
    bsp_src = open(os.path.join(REF, "elsa/projectors/tests/test_BSplines.cpp")).read()
    at = bsp_src.index("if (derivative) {", bsp_src.index("getExpected(int order, int dim, bool derivative)\n{"))
    bsp_d = numbers(brace_block(bsp_src, bsp_src.index("return", at)))
    assert len(bsp_d) == 101, len(bsp_d)
    json.dump({
        "blob_derivative": {"source": "elsa/projectors/tests/test_BlobDerivative.cpp", "a": 1.0,
                            "s": "i / 100", "epsilon": "doctest Approx default",
                            "gradient_helper_same_numbers": same,
                            "tables": [{"alpha": k[0], "m": k[1], "expected": v} for k, v in sorted(deriv.items())]},
        "bspline_derivative_lut": {"source": "elsa/projectors/tests/test_BSplines.cpp", "order": 3, "dim": 2,
                                   "N": 101, "x": "i / 50", "expected": bsp_d},
    }, open(os.path.join(OUT, "derivatives.json"), "w"), indent=1)
    bvp_src = open(os.path.join(REF, "elsa/projectors/tests/test_BlobVoxelProjector.cpp")).read()
    at = bvp_src.index('TEST_CASE_TEMPLATE("BlobVoxelProjector: Test weights"')
    weights = numbers(brace_block(bvp_src, bvp_src.index("expected", at)))
    assert len(weights) == 51, len(weights)
    json.dump({"source": "elsa/projectors/tests/test_BlobVoxelProjector.cpp", "radius": 2.0, "alpha": 10.83,
               "order": 2, "distance": "i / 25", "epsilon": "doctest Approx default", "expected": weights},
              open(os.path.join(OUT, "blob_weights.json"), "w"), indent=1)
    print("wrote", os.listdir(OUT))


if __name__ == "__main__":
    sys.exit(main())
