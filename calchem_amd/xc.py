"""Exchange-correlation functional names and ids (crates/calchem-dft/src/xc.rs:16-100)."""
from __future__ import annotations

XC_IDS = {"svwn5": 0, "blyp": 1, "pbe": 2, "b3lyp": 3, "b3lyp5": 4, "pbe0": 5}

_ALIASES = {
    "svwn": "svwn5", "svwn5": "svwn5", "lda": "svwn5", "slatervwn5": "svwn5",
    "blyp": "blyp", "pbe": "pbe", "pbepbe": "pbe", "b3lyp": "b3lyp", "b3lyp5": "b3lyp5",
    "pbe0": "pbe0", "pbe1pbe": "pbe0",
}

# Exact-exchange fractions (xc.rs:116-128); the grid part excludes them.
EXACT_EXCHANGE = {"svwn5": 0.0, "blyp": 0.0, "pbe": 0.0, "b3lyp": 0.20, "b3lyp5": 0.20, "pbe0": 0.25}


def canonical(name: str) -> str:
    """XcFunctional::from_name normalisation: lower case, drop blanks, '-' and '_'."""
    key = "".join(ch for ch in name.strip().lower() if not ch.isspace() and ch not in "-_")
    if key not in _ALIASES:
        raise ValueError(f"unsupported DFT XC functional {name!r} (supported: {sorted(XC_IDS)})")
    return _ALIASES[key]


def xc_id(name: str) -> int:
    return XC_IDS[canonical(name)]
