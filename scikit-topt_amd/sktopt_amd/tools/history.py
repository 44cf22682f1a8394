"""Per-iteration records and their ``histories.npz`` file.

Call surface and file format of ``tools/history.py`` (``HistorySeries``
:62-230, ``HistoryCollection`` :233-700): a history registered with
``plot_type="value"`` keeps one float per iteration; with
``"min-max-mean"`` / ``"min-max-mean-std"`` an array input is reduced to
``[min, mean, max(, std)]`` (``:125-154``).  ``export_histories`` writes
``{name: data, f"{name}_header": [name, plot_type]}`` with ``numpy.savez``
(``:536-610``), where aggregated data is stored transposed,
``(n_stats, n_iters)`` (``:172-206``); ``import_histories`` reads it back
(``:612-700``).  Plots (``export_progress``) are not produced.

One extension: ``plot_type="raw"`` keeps each fed value as given (used for
the per-load-case compliances), exported as a stacked array.
"""
from __future__ import annotations

import logging
import os

import numpy as np

logger = logging.getLogger(__name__)

_STATS = ("min-max-mean", "min-max-mean-std")


class HistorySeries:
    def __init__(self, name, constants=None, constant_names=None, plot_type="value",
                 ylog=False, data=None):
        self.name = name
        self.constants = constants
        self.constant_names = constant_names
        self.plot_type = plot_type
        self.ylog = ylog
        self.data = list(data) if data is not None else []

    def exists(self):
        return len(self.data) > 0

    def add(self, data_input):
        if self.plot_type == "raw":
            self.data.append(np.array(data_input, dtype=float, copy=True))
            return
        if isinstance(data_input, np.ndarray):
            if data_input.shape == ():
                self.data.append(float(data_input))
            else:
                tmp = [float(np.min(data_input)), float(np.mean(data_input)),
                       float(np.max(data_input))]
                if self.plot_type == "min-max-mean-std":
                    tmp.append(float(np.std(data_input)))
                self.data.append(tmp)
        else:
            self.data.append(float(data_input))

    def add_stats(self, stats):
        """Append precomputed ``[min, mean, max(, std)]`` (statistics taken
        on the device); the same entry ``add`` makes from the full array."""
        n = 4 if self.plot_type == "min-max-mean-std" else 3
        self.data.append([float(v) for v in list(stats)[:n]])

    def latest(self):
        d = self.data[-1]
        return np.array(d) if isinstance(d, list) else d

    def data_to_array(self):
        header = [self.name, self.plot_type]
        if len(self.data) == 0:
            return np.array([]), header
        if self.plot_type == "raw":
            return np.stack(self.data), header
        if isinstance(self.data[0], list):
            data = np.array(self.data)
            data = data.T[:4] if self.plot_type == "min-max-mean-std" else data.T[:3]
        else:
            data = np.array(self.data)
        return data, header


class HistoryCollection:
    def __init__(self, dst_path: str | None = None):
        self.dst_path = dst_path
        self.histories: dict[str, HistorySeries] = {}

    def add(self, name, constants=None, constant_names=None, plot_type="value", ylog=False,
            data=None):
        self.histories[name] = HistorySeries(name, constants, constant_names, plot_type, ylog,
                                             data)

    def feed_data(self, name, data):
        if name not in self.histories:           # records this build adds
            self.add(name, plot_type="raw" if np.ndim(data) > 0 else "value")
        self.histories[name].add(data)

    def feed_stats(self, name, stats):
        self.histories[name].add_stats(stats)

    def latest(self, name):
        h = self.histories.get(name)
        return h.latest() if h is not None and h.exists() else None

    @property
    def data(self):
        return {k: v.data for k, v in self.histories.items()}

    def as_object(self):
        return type("Histories", (), {k: list(v.data) for k, v in self.histories.items()})

    def as_object_latest(self):
        return type("Latest", (), {k: (v.latest() if v.exists() else None)
                                   for k, v in self.histories.items()})

    def histories_to_array(self):
        out = {}
        for name, h in self.histories.items():
            if not h.exists():
                continue
            data, header = h.data_to_array()
            out[name] = data
            out[f"{name}_header"] = np.array(header, dtype=str)
        return out

    def export_histories(self, fname: str | None = None):
        fname = fname or "histories.npz"
        arrays = self.histories_to_array()
        if not any(not k.endswith("_header") for k in arrays):
            logger.warning("No histories to save.")
            return
        if not isinstance(self.dst_path, str):
            logger.warning("Invalid destination path.")
            return
        np.savez(os.path.join(self.dst_path, fname), **arrays)

    def import_histories(self, fname: str | None = None):
        fname = fname or "histories.npz"
        path = os.path.join(self.dst_path, fname) if isinstance(self.dst_path, str) else None
        if path is None or not os.path.exists(path):
            logger.warning(f"File not found: {path}")
            return
        new = {}
        with np.load(path, allow_pickle=False) as d:
            for key in d.files:
                if key.endswith("_header"):
                    continue
                arr = d[key]
                hk = f"{key}_header"
                if hk in d.files:
                    header = d[hk].tolist()
                    name, plot_type = header[0], (header[1] if len(header) > 1
                                                  else "min-max-mean")
                else:
                    name, plot_type = key, "value"
                old = self.histories.get(name)
                h = HistorySeries(name, old.constants if old else None,
                                  old.constant_names if old else None, plot_type,
                                  old.ylog if old else False)
                if plot_type == "raw":
                    h.data = [np.array(r) for r in arr]
                elif arr.ndim == 2 and plot_type in _STATS:
                    h.data = [list(map(float, c)) for c in arr.T]
                else:
                    h.data = [float(v) for v in np.ravel(arr)]
                new[name] = h
        self.histories = new
