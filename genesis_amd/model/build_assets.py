"""Regenerate the saved loader outputs under genesis_amd/assets/specs/ from MJCF sources.

Usage: python -m genesis_amd.model.build_assets [asset_root]
The Franka Panda MJCF and meshes are read from ``asset_root`` (default: the reference checkout's asset directory);
only the loader's output (link/joint/geom parameters and convex-hull vertices) is saved.
"""

from __future__ import annotations

import os
import sys

from ..morphs import cached_spec_path
from .mjcf import load_mjcf
from .spec_io import save_spec

ASSETS = ["xml/franka_emika_panda/panda.xml"]


def main(asset_root: str = "/root/reference/genesis/assets") -> None:
    for rel in ASSETS:
        src = os.path.join(asset_root, rel)
        spec = load_mjcf(src)
        out = cached_spec_path(rel)
        os.makedirs(os.path.dirname(out), exist_ok=True)
        save_spec(spec, out)
        print(f"{rel} -> {out} ({os.path.getsize(out)} bytes)")


if __name__ == "__main__":
    main(*sys.argv[1:])
