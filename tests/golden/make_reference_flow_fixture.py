"""Writes tests/golden/reference_flow.npz: the reference's trained flow
(assets/rq_spline_model.bson — Dense 24→768 ReLU, Dense 768→34, K = 10, the
eight label normalisers, the hyper-parameters), read with the package's
data-only BSON reader.  Run in a checkout next to the reference:

    python tests/golden/make_reference_flow_fixture.py /path/to/reference/assets/rq_spline_model.bson
"""
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from neutrinotelescopes_amd.refflow import load_reference_flow  # noqa: E402

if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/assets/rq_spline_model.bson"
    model = load_reference_flow(src)
    out = Path(__file__).resolve().parent / "reference_flow.npz"
    model.to_npz(out)
    print(f"wrote {out}: W1 {model.w1.shape}, W2 {model.w2.shape}, K = {model.n_bins}")
