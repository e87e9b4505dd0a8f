"""Writes the committed workload fixtures of tests/golden/ from
highmap_amd.workloads (run from the repository root):

    python tests/golden/make_fixtures.py

batch_manifest.json      the 96-map batching workload (seeds 100..195)
serve_requests_head.json the first 50 requests of the serving stream
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from highmap_amd import workloads  # noqa: E402


def main():
    with open(os.path.join(HERE, "batch_manifest.json"), "w") as f:
        json.dump(workloads.batch_manifest(), f, indent=1)
    with open(os.path.join(HERE, "serve_requests_head.json"), "w") as f:
        json.dump(workloads.serve_requests()[:50], f, indent=1)


if __name__ == "__main__":
    main()
