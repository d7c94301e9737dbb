"""Golden fixtures: inputs and outputs of reference routines, with the scripts
that wrote them."""
from pathlib import Path

HERE = Path(__file__).resolve().parent


def golden_path(name: str) -> Path:
    return HERE / name
