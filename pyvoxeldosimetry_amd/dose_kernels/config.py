"""Nuclide configuration files in the reference's JSON layout.

``BaseKernelGenerator._load_config`` (``pyvoxeldosimetry/data/dose_kernels/base_kernel.py:34-36``)
reads ``<Nuclide>/<Nuclide>.json`` with ``json.load``.  Five of the reference's six
files carry ``//`` comments (``Lu177/Lu177.json:16``, ``Ga68/Ga68.json:16`` ...), which
``json.load`` rejects; ``load_config`` strips them first (outside string literals),
so both those files and this package's own (``dose_kernels/data``) load.

Fields read (reference layout): ``nuclide.{name, symbol, half_life}``,
``nuclide.particle_energies.{beta-_max, beta-_abundance, beta+_max, annihilation,
gamma_lines[{energy, intensity}], bremsstrahlung}``, ``kernel.{grid_size,
scaling_factor}``.
"""
from __future__ import annotations

import json
import os
from typing import Union

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def strip_json_comments(text: str) -> str:
    """Remove ``// ...`` comments that lie outside string literals (escapes honoured)."""
    out = []
    i, n = 0, len(text)
    in_str = False
    while i < n:
        c = text[i]
        if in_str:
            out.append(c)
            if c == "\\" and i + 1 < n:
                out.append(text[i + 1])
                i += 2
                continue
            if c == '"':
                in_str = False
            i += 1
            continue
        if c == '"':
            in_str = True
            out.append(c)
            i += 1
            continue
        if c == "/" and i + 1 < n and text[i + 1] == "/":
            j = text.find("\n", i)
            i = n if j < 0 else j
            continue
        out.append(c)
        i += 1
    return "".join(out)


def load_config(path: Union[str, os.PathLike]) -> dict:
    """Parse a nuclide JSON file of the reference's layout, ``//`` comments allowed."""
    with open(path, "r") as f:
        return json.loads(strip_json_comments(f.read()))


def packaged_config(symbol: str) -> dict:
    """This package's own file for ``symbol`` (same layout, no comments)."""
    return load_config(os.path.join(DATA_DIR, f"{symbol}.json"))
