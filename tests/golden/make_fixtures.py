"""Writes tests/golden/coffin_frames.npz from the reference's recorded frames.

Reads data/depth_{0,6}.bin and data/pose_{0,6}.bin of the reference
repository (Eigen ``read_binary`` layout, src/testMRFMapViewer.cpp:18-27:
int64 rows, int64 cols, then rows*cols float32 column-major) and stores them
row-major together with cam_in_body of config/coffin_world_640.yaml:2-5.
These are input fixtures only: the reference publishes no outputs for them.
Run in a checkout where the reference is mounted (default /root/reference).
"""
import os
import sys

import numpy as np
import yaml

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def read_binary(path):
    with open(path, "rb") as f:
        rows, cols = np.frombuffer(f.read(16), dtype=np.int64)
        data = np.frombuffer(f.read(int(rows) * int(cols) * 4), dtype=np.float32)
    return data.reshape(int(cols), int(rows)).T.copy()


def main():
    frames = [0, 6]
    depth = np.stack([read_binary(os.path.join(REF, "data", f"depth_{k}.bin")) for k in frames])
    pose = np.stack([read_binary(os.path.join(REF, "data", f"pose_{k}.bin")) for k in frames])
    with open(os.path.join(REF, "config", "coffin_world_640.yaml")) as f:
        cfg = yaml.safe_load(f)
    cam_in_body = np.asarray(cfg["viewer_params_nodes"]["cam_in_body"], dtype=np.float32).reshape(4, 4)
    np.savez_compressed(os.path.join(HERE, "coffin_frames.npz"), frames=np.asarray(frames), depth=depth,
                        body_pose=pose, cam_in_body=cam_in_body)


if __name__ == "__main__":
    main()
