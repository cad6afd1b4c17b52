"""Writes cd2_fel_mle.json: the per-site FEL table of the reference's
HyPhy 2.5 ground truth for CD2 (benchmarks/ground_truth/cd2_fel.json, the
"MLE" content: alpha, beta, alpha=beta, LRT, p-value, total branch length
per codon site) and the global MG94 fit the reference's matched-tree parity
test starts from (tests/test_fel_cd2_parity.cpp:67-80: exchangeabilities and
omega; the tree is benchmarks/data/cd2_fitted.nwk, kept beside this file).

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_fel_truth.py
"""
import json
import os

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    d = json.load(open(os.path.join(REF, "benchmarks", "ground_truth", "cd2_fel.json")))
    out = {
        "source": "benchmarks/ground_truth/cd2_fel.json (HyPhy 2.5.101 FEL on benchmarks/data/cd2.fna)",
        "headers": [h[0] for h in d["MLE"]["headers"]],
        "mle": d["MLE"]["content"]["0"],
        "global_mg94": {
            "source": "tests/test_fel_cd2_parity.cpp:67-80",
            "theta": [0.5521777318173139, 0.2653120245806623, 0.4932059218537349, 1.03133955537726,
                      0.3054313963126315],
            "alpha": 1.0,
            "beta": 1.012387352492807,
            "log_likelihood": d["fits"]["Global MG94xREV"]["Log Likelihood"],
        },
        "thresholds": {"source": "tests/test_fel_cd2_parity.cpp:117-120", "lrt_pearson": 0.998,
                       "lrt_spearman": 0.995, "p_pearson": 0.993, "p_spearman": 0.995},
    }
    with open(os.path.join(HERE, "cd2_fel_mle.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
