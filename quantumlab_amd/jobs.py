"""QuantumLab's job configuration files (Bin/*.yaml) and the parameters they
carry into this library's calls.

The applications read these files with CYAMLParser (QL/ToolClass/
CYAMLParser.h:70-223): a YAML subset of `Key : value` lines, `[a, b, c]`
lists, `Key:` lines opening an indented block of further keys, and `#`
comments; every value is kept as a string and converted when fetched
(CParameters::FetchValueINT / Real / ...). `read_job` reads the same subset into
nested dicts of strings and lists of strings; `Job` fetches typed values.

`swap_test_job` maps Bin/SwapTest.yaml (read by APP/SwapTest/SwapTest.cpp) to
what this library takes: the noise-after-gate rates as a qmxNoiseAfterGate
(QLSimulatorDensityMatrix.cpp:88-153 applies them after every gate) and the
before-measure rates, which QLSimulatorDensityMatrix.cpp:190-216 applies to
every qubit before each measurement (ops appended to an operation list here).
"""
from __future__ import annotations

import dataclasses
import re
from typing import Dict, List, Union

Value = Union[str, List[str], "Dict[str, Value]"]


def _strip_comment(line: str) -> str:
    i = line.find("#")
    return line if i < 0 else line[:i]


def read_job(text: str) -> Dict[str, Value]:
    """Parse the YAML subset QuantumLab's job files use."""
    root: Dict[str, Value] = {}
    stack = [(-1, root)]  # (indent of the block's keys' parent, dict)
    for raw in text.splitlines():
        line = _strip_comment(raw).rstrip()
        if not line.strip():
            continue
        indent = len(line) - len(line.lstrip(" \t"))
        body = line.strip()
        if ":" not in body:
            raise ValueError(f"not a 'key : value' line: {raw!r}")
        key, value = body.split(":", 1)
        key, value = key.strip(), value.strip()
        if not key:
            raise ValueError(f"empty key: {raw!r}")
        while len(stack) > 1 and indent <= stack[-1][0]:
            stack.pop()
        parent = stack[-1][1]
        if value == "":
            block: Dict[str, Value] = {}
            parent[key] = block
            stack.append((indent, block))
        elif value.startswith("["):
            if not value.endswith("]"):
                raise ValueError(f"unterminated list: {raw!r}")
            inner = value[1:-1].strip()
            parent[key] = [v.strip() for v in inner.split(",")] if inner else []
        else:
            parent[key] = value
    return root


def read_job_file(path: str) -> Dict[str, Value]:
    with open(path, "r", encoding="utf-8", errors="replace") as f:
        return read_job(f.read())


@dataclasses.dataclass
class Job:
    """Typed access to one block, as CParameters::Fetch* converts strings."""
    values: Dict[str, Value]

    def block(self, key: str) -> "Job":
        v = self.values[key]
        if not isinstance(v, dict):
            raise KeyError(f"{key} is not a block")
        return Job(v)

    def string(self, key: str, default: str | None = None) -> str:
        v = self.values.get(key, default)
        if v is None or not isinstance(v, str):
            raise KeyError(key)
        return v

    def int(self, key: str, default: int | None = None) -> int:
        if key not in self.values and default is not None:
            return default
        return int(float(self.string(key)))

    def real(self, key: str, default: float | None = None) -> float:
        if key not in self.values and default is not None:
            return default
        return float(self.string(key))

    def reals(self, key: str) -> List[float]:
        v = self.values[key]
        return [float(x) for x in (v if isinstance(v, list) else [v])]

    def ints(self, key: str) -> List[int]:
        return [int(float(x)) for x in self.reals(key)]


@dataclasses.dataclass
class SwapTestJob:
    """Bin/SwapTest.yaml as APP/SwapTest/SwapTest.cpp reads it."""
    dot_type: str                 # EDT_AmplitudeEncodeComplex / EDT_AmplitudeEncodeReal / EDT_SimpleEncode
    measure_repeat: int
    enable_noise: bool
    enable_classical: bool
    simple_encode_qubits: int
    noise_after_gate: Dict[str, float]    # fields of qmxNoiseAfterGate
    before_measure: Dict[str, float]      # damping, depolarising, dephasing, mixPauli on every qubit
    files: Dict[str, str]


_NOISE_KEYS = {"damping": "DampingAfterGate", "depolarising": "DepolarisingAfterGate",
               "dephasing": "DephasingAfterGate", "mixPauli": "MixPauliAfterGate",
               "twoQubitDephasing": "TwoQubitDephasing", "twoQubitDepolarising": "TwoQubitDepolarising"}
_BEFORE_KEYS = {"damping": "DampingBeforeMeasure", "depolarising": "DepolarisingBeforeMeasure",
                "dephasing": "DephasingBeforeMeasure", "mixPauli": "MixPauliBeforeMeasure"}


def swap_test_job(values: Dict[str, Value]) -> SwapTestJob:
    j = Job(values)
    files = {k: j.string(k) for k in ("FileName1", "FileName2", "ClassicalResSaveFileName",
                                      "QuantumResSaveFileName", "AveragePuritySaveFileName",
                                      "AverageFidelitySaveFileName") if k in values}
    return SwapTestJob(dot_type=j.string("DotType"), measure_repeat=j.int("MeasureRepeat", 0),
                       enable_noise=j.int("EnableNoise", 0) != 0, enable_classical=j.int("EnableClassical", 0) != 0,
                       simple_encode_qubits=j.int("SimpleEncodeQubits", 0),
                       noise_after_gate={f: j.real(k, 0.0) for f, k in _NOISE_KEYS.items()},
                       before_measure={f: j.real(k, 0.0) for f, k in _BEFORE_KEYS.items()}, files=files)


def noise_struct(qe, job: SwapTestJob):
    """The job's after-gate rates as the binding's qmxNoiseAfterGate."""
    return qe.t.NoiseAfterGate(**job.noise_after_gate)


_NAME = re.compile(r"^[A-Za-z_][A-Za-z0-9_]*$")


def is_enum_name(value: str) -> bool:
    """Job values such as `EDT_SimpleEncode` name enumerators of the app."""
    return bool(_NAME.match(value))
