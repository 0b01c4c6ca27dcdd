"""Copy the reference's cubic-kernel golden vectors into a fixture.

Source (data file held by the reference's own tests):
  tools/integrity/integrity/cat3_numerical/expected_values.json
(6 points at h = 1, atol = rtol = 1e-5; generated in closed form by the
reference's generate_expected.py from SPlisHSPlasH 2.16.1 SPHKernels.h).
Run once where the reference tree is mounted:
  python tests/golden/make_cubic_kernel_fixture.py /root/reference
"""
import json
import os
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = os.path.join(ref, "tools/integrity/integrity/cat3_numerical/expected_values.json")
with open(src) as f:
    data = json.load(f)
out = {
    "source": "tools/integrity/integrity/cat3_numerical/expected_values.json",
    "tolerance": data["tolerance"],
    "points": [{"q": p["q"], "h": p["h"], "W": p["expected_W"],
                "gradW_magnitude": p["expected_gradW_magnitude"]} for p in data["test_points"]],
}
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cubic_kernel_expected.json")
with open(dst, "w") as f:
    json.dump(out, f, indent=1)
print(f"wrote {dst} ({len(out['points'])} points)")
