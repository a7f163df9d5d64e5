"""Extract golden vectors that the reference holds as literal data (read as text, never imported).

* fibo45: the 45 precomputed seed quaternions of _get_45_fibo
  (shepherd_score/alignment/_torch.py:199-250) -- the expected output of the Fibonacci-sphere
  seed construction for 45 points.
* se3_known_answer: the quaternion / SE(3) / transformed-points known answers of
  tests/test_alignment_utils.py:44-71 (values from an external rotation converter).

Writes tests/golden/reference_tables.json.  Run in a container that has the reference mounted:
    python tests/golden/make_fixtures.py /root/reference
"""
import ast
import json
import os
import re
import sys


def _array_after(text: str, anchor: str) -> list:
    i = text.index(anchor)
    j = text.index("[", i)
    depth, k = 0, j
    while True:
        if text[k] == "[":
            depth += 1
        elif text[k] == "]":
            depth -= 1
            if depth == 0:
                break
        k += 1
    return ast.literal_eval(re.sub(r"\s+", " ", text[j:k + 1]))


def main(ref_root: str) -> None:
    torch_src = open(os.path.join(ref_root, "shepherd_score", "alignment", "_torch.py")).read()
    fibo = _array_after(torch_src, "def _get_45_fibo")
    test_src = open(os.path.join(ref_root, "tests", "test_alignment_utils.py")).read()
    out = {
        "fibo45": fibo,
        "se3_known_answer": {
            "quaternion": _array_after(test_src, "ex_quaternion = np.array("),
            "rotation": _array_after(test_src, "sol_rot_matrix = np.array("),
            "se3_params": _array_after(test_src, "ex_se3_params = np.array("),
            "se3_transform": _array_after(test_src, "sol_se3_transform = np.array("),
            "points": _array_after(test_src, "ex_set_of_points = np.array("),
        },
        "source": {
            "fibo45": "shepherd_score/alignment/_torch.py:199-250",
            "se3_known_answer": "tests/test_alignment_utils.py:44-71",
        },
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tables.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst, len(fibo), "fibo rows")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
