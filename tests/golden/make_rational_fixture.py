"""Extracts a few x^(-1/2) entries (order, precision, interval) of the
reference's rational-approximation table (include/approx.h) into
rational_table_sample.json.  Run in the build container, where the
reference is mounted; the fixture is committed so no test reads the
reference at run time."""
import json
import os
import re

SRC = "/root/reference/include/approx.h"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rational_table_sample.json")


def main():
    text = open(SRC).read()
    body = text[text.index("{") + 1: text.rindex("}")]
    nums = [float(v) for v in re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", body)]
    entries = []
    i = 0
    while i + 7 <= len(nums):
        n, d, order, prec, lo, hi = nums[i:i + 6]
        order = int(order)
        if int(n) == 1 and int(d) == 2 and order in (10, 12, 16, 20, 24):
            entries.append({"n": -1, "d": 2, "order": order, "prec": prec, "min": lo, "max": hi})
        i += 7 + 2 * order
    json.dump({"source": "include/approx.h (order, precision, interval of x^(-1/2) entries)",
               "entries": entries[:12]}, open(OUT, "w"), indent=1)
    print(len(entries), "entries matched;", min(len(entries), 12), "written")


if __name__ == "__main__":
    main()
