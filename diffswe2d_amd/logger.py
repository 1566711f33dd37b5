"""Run logger of the example drivers (src/logger.py:3-23): terminal lines without
timestamps, and the same lines with timestamps in a log file opened for writing."""
from __future__ import annotations

import logging


def setup_logger(log_filename="simulation.log"):
    log = logging.getLogger("DiffSWE")
    log.setLevel(logging.INFO)
    log.handlers.clear()
    terminal = logging.StreamHandler()
    terminal.setFormatter(logging.Formatter("%(message)s"))
    to_file = logging.FileHandler(log_filename, mode="w")
    to_file.setFormatter(logging.Formatter("%(asctime)s - %(message)s", datefmt="%Y-%m-%d %H:%M:%S"))
    log.addHandler(terminal)
    log.addHandler(to_file)
    return log
