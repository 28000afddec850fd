"""Regenerates the golden fixtures under tests/golden/data from the reference.

Needs the reference at /root/reference and the Node binary on PATH; run from the
repository root:  python tests/golden/make_fixtures.py [job ...]
Writes each scene's JSON (bearing_amd/scenes.py), then runs make_fixtures.mjs under
the lower.mjs loader once per job. The fixtures are what the tests read; this
script is not run by any test.
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from bearing_amd import scenes  # noqa: E402

DATA = os.path.join(HERE, "data")

# name: scene, dt, steps, soft-phase steps, colour steps, trace step, transforms every
JOBS = {
    "pyramid": dict(scene="pyramid", dt=1 / 60, steps=600, soft=[1, 2, 30, 55, 63], colour=[1, 30, 62, 66], trace=3,
                    every=100),
    "rain_small": dict(scene="rain_small", dt=1 / 60, steps=240, soft=[20, 40, 60, 90], colour=[60], trace=45, every=60),
    "tower": dict(scene="tower", dt=1 / 60, steps=120, soft=[1, 10, 40, 64], colour=[10, 63], trace=5, every=30),
    "throw": dict(scene="throw", dt=1 / 120, steps=120, soft=[30, 60, 90, 119], colour=[60], trace=70, every=30),
    "pile_small": dict(scene="pile_small", dt=1 / 60, steps=150, soft=[30, 70, 110], colour=[70], trace=100, every=50),
    # The 2,000-sphere rain at full size: its solve phase mid-landing is the bench's workload.
    "rain": dict(scene="rain", dt=1 / 60, steps=900, soft=[300], colour=[], trace=-1, every=300),
}


def run(name):
    job = dict(JOBS[name])
    os.makedirs(os.path.join(DATA, "scenes"), exist_ok=True)
    scene_path = os.path.join(DATA, "scenes", job["scene"] + ".json")
    with open(scene_path, "w") as f:
        json.dump(scenes.SCENES[job["scene"]](), f)
    job["scene"] = scene_path
    job["out"] = os.path.join(DATA, name)
    job_path = os.path.join(DATA, name + ".job.json")
    with open(job_path, "w") as f:
        json.dump(job, f)
    subprocess.run(["node", "--no-warnings", "--experimental-loader", os.path.join(HERE, "lower.mjs"),
                    os.path.join(HERE, "make_fixtures.mjs"), job_path], check=True)


if __name__ == "__main__":
    for name in sys.argv[1:] or [j for j in JOBS if j != "rain"]:
        run(name)
