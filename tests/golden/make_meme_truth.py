"""Writes cd2_meme_mle.json: the per-site MEME table of the reference's
HyPhy 2.5 ground truth for CD2 (benchmarks/ground_truth/cd2_meme_rel.json,
the "MLE" content: alpha, beta1, p1, beta+, p+, LRT, p-value, branches under
selection, total branch length, MEME lnL, FEL lnL, MEME-vs-FEL p-value, FEL
alpha, FEL beta per codon site) and the thresholds of the reference's MEME
parity test (tests/test_meme.cpp:66-103).

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_meme_truth.py
"""
import json
import os

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    d = json.load(open(os.path.join(REF, "benchmarks", "ground_truth", "cd2_meme_rel.json")))
    out = {
        "source": "benchmarks/ground_truth/cd2_meme_rel.json (HyPhy 2.5 MEME on benchmarks/data/cd2.fna)",
        "headers": [h[0] for h in d["MLE"]["headers"]],
        "mle": d["MLE"]["content"]["0"],
        "global_mg94_log_likelihood": d["fits"]["Global MG94xREV"]["Log Likelihood"],
        "thresholds": {"source": "tests/test_meme.cpp:98-102", "lrt_pearson": 0.95, "lrt_spearman": 0.95,
                       "p_pearson": 0.95, "p_spearman": 0.95},
    }
    with open(os.path.join(HERE, "cd2_meme_mle.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
