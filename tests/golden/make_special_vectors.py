"""Writes tests/golden/special_vectors.json: 50-digit mpmath values of the
special functions the fit and test_de use (digamma, trigamma, log B(a,b) and
the two-sided Student-t tail 2*P(T_df > t) = I_{df/(df+t^2)}(df/2, 1/2)).

These are the fixed points the oracle's special functions are pinned to.
Run:  python tests/golden/make_special_vectors.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    xs = [1e-10, 1e-6, 1e-3, 0.1, 0.5, 1.0, 1.4616321449683622, 2.0, 3.7, 9.5, 10.0,
          10.5, 33.3, 100.0, 1e3, 1e5, 1e8, 1e12]
    out = {"digamma": [], "trigamma": [], "lbeta": [], "pt2_log": []}
    for x in xs:
        out["digamma"].append([x, mp.nstr(mp.digamma(x), 30)])
        out["trigamma"].append([x, mp.nstr(mp.polygamma(1, x), 30)])
    for a, b in [(0.5, 0.5), (0.5, 3.0), (2.5, 0.5), (12.0, 0.5), (11.0, 13.0),
                 (2499.0, 0.5), (49999.0, 0.5), (249999.0, 0.5), (1e6, 0.5)]:
        out["lbeta"].append([a, b, mp.nstr(mp.log(mp.beta(a, b)), 30)])
    for df in [3.0, 30.0, 4998.0, 99998.0, 499998.0]:
        for t in [0.05, 0.5, 1.0, 1.96, 3.0, 6.0, 12.0, 40.0]:
            x = mp.mpf(df) / (df + mp.mpf(t) ** 2)
            v = mp.betainc(mp.mpf(df) / 2, mp.mpf(0.5), 0, x, regularized=True)
            out["pt2_log"].append([t, df, mp.nstr(mp.log(v), 30)])
    with open(os.path.join(HERE, "special_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
