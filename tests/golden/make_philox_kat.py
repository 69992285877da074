"""Write tests/golden/philox4x32_10_kat.json from the Random123 known-answer
file that RandBLAS ships (test/test_basic_rng/r123_kat_vectors.txt:19-21).

Run in the build container only (it reads /root/reference); the JSON it
writes is the committed fixture: data rows of (counter, key, expected words).
"""
import json
import os

SRC = "/root/reference/test/test_basic_rng/r123_kat_vectors.txt"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "philox4x32_10_kat.json")

rows = []
with open(SRC) as f:
    for line in f:
        t = line.split()
        if len(t) >= 12 and t[0] == "philox4x32" and t[1] == "10":
            w = [int(x, 16) for x in t[2:12]]
            rows.append({"counter": w[0:4], "key": w[4:6], "expected": w[6:10]})
with open(OUT, "w") as f:
    json.dump({"source": "test/test_basic_rng/r123_kat_vectors.txt:19-21 (philox4x32, 10 rounds)",
               "rows": rows}, f, indent=1)
print(f"wrote {len(rows)} rows to {OUT}")
