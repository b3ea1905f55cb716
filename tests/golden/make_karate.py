"""Writes karate.json: the Zachary karate club graph and its triangle count.

The edge arrays are read as data from the LAGraph test header
baselines/LAGraph/src/test/include/graph_zachary_karate.h:25-99 (ZACHARY_I,
ZACHARY_J; 34 nodes, 156 directed entries); the expected count 45 is the
known answer in baselines/LAGraph/src/test/test_TriangleCount.c:38-51.
Run once in a tree that holds the reference; the fixture is committed.
"""
import json
import os
import re
import sys

REF = os.environ.get("BTC_REFERENCE", "/root/reference")
HDR = os.path.join(REF, "baselines/LAGraph/src/test/include/graph_zachary_karate.h")


def array(text, name):
    m = re.search(name + r"\[\]\s*=\s*\{([^}]*)\}", text)
    return [int(x) for x in re.findall(r"\d+", m.group(1))]


def main():
    text = open(HDR).read()
    i = array(text, "ZACHARY_I")
    j = array(text, "ZACHARY_J")
    assert len(i) == len(j) == 156, (len(i), len(j))
    out = {
        "name": "zachary_karate",
        "source": "baselines/LAGraph/src/test/include/graph_zachary_karate.h (ZACHARY_I/J)",
        "expected_source": "baselines/LAGraph/src/test/test_TriangleCount.c:38-51",
        "n": 34,
        "src": i,
        "dst": j,
        "triangles": 45,
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "karate.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print(path, len(i), "entries")


if __name__ == "__main__":
    sys.exit(main())
