"""Package logger (same name and format as the reference's util/logger.py:22-41)."""

import logging
import sys

logger = logging.getLogger("fbpinns")
if not logger.handlers:
    _h = logging.StreamHandler(sys.stdout)
    _h.setFormatter(logging.Formatter("[%(levelname)s] %(asctime)s - %(message)s"))
    logger.addHandler(_h)
    logger.setLevel(logging.INFO)
    logger.propagate = False


def switch_to_file_logger(filename):
    """Route the package log to a file (util/logger.py:9-16)."""
    for h in list(logger.handlers):
        logger.removeHandler(h)
    h = logging.FileHandler(filename)
    h.setFormatter(logging.Formatter("[%(levelname)s] %(asctime)s - %(message)s"))
    logger.addHandler(h)
