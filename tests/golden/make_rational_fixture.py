"""Extracts entries of the reference's rational-approximation table
(include/approx.h) into rational_table_sample.json: (order, precision,
interval) of x^(-1/2) and x^(-1/4) entries, and the full coefficients
(a0, roots, poles) of one entry of each power.  Run in the build container,
where the reference is mounted; the fixture is committed so no test reads
the reference at run time.

Table layout (library/rational.cpp:64-107): per entry n, d, order, prec,
min, max, a0, then `order` roots and `order` poles; the entries approximate
x^(-n/d) (rational_find swaps roots and poles for positive powers)."""
import json
import os
import re

SRC = "/root/reference/include/approx.h"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rational_table_sample.json")
ORDERS = (10, 12, 16, 20, 24, 30)


def parse():
    text = open(SRC).read()
    body = text[text.index("{") + 1: text.rindex("}")]
    nums = [float(v) for v in re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", body)]
    entries = []
    i = 0
    while i + 7 <= len(nums):
        n, d, order, prec, lo, hi, a0 = nums[i:i + 7]
        order = int(order)
        roots = nums[i + 7: i + 7 + order]
        poles = nums[i + 7 + order: i + 7 + 2 * order]
        entries.append({"n": -int(n), "d": int(d), "order": order, "prec": prec, "min": lo, "max": hi,
                        "a0": a0, "root": roots, "pole": poles})
        i += 7 + 2 * order
    return entries


def main():
    entries = parse()
    sample = [{k: e[k] for k in ("n", "d", "order", "prec", "min", "max")}
              for e in entries if e["order"] in ORDERS]
    coeffs = [e for e in entries if e["order"] == 12 and e["prec"] < 1e-10 and e["prec"] > 9e-11]
    json.dump({"source": "include/approx.h: (order, precision, interval) of x^(-1/2) and x^(-1/4) entries, "
                         "and the coefficients of the order-12, 1e-10 entries",
               "entries": sample, "coefficients": coeffs}, open(OUT, "w"), indent=1)
    print(len(entries), "table entries;", len(sample), "sampled;", len(coeffs), "with coefficients")


if __name__ == "__main__":
    main()
