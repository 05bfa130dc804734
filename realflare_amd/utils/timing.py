"""Stage timing in the reference's log format (``realflare/utils/timing.py:8-30``):
``f'{qualname + ":": <40}{ms:9.3f}ms'`` at INFO.

The stages enqueue device work asynchronously.  When INFO is enabled the
device is synchronized before the clock starts and before it stops, so a
logged time is the stage's host work plus the device work it enqueued (as the
reference's blocking OpenCL calls measure); with INFO off nothing is
synchronized and nothing is logged.
"""
import logging
import time
from functools import wraps

logger = logging.getLogger(__name__)


def _sync() -> None:
    import torch

    if torch.cuda.is_available() and torch.cuda.is_initialized():
        torch.cuda.synchronize()


def timer(func):
    @wraps(func)
    def wrapper(*args, **kwargs):
        if not logger.isEnabledFor(logging.INFO):
            return func(*args, **kwargs)
        _sync()
        start = time.perf_counter()
        result = func(*args, **kwargs)
        _sync()
        label = f'{func.__qualname__}:'
        logger.info(f'{label: <40}{(time.perf_counter() - start) * 1000:9.3f}ms')
        return result

    return wrapper


class Timer:
    def __init__(self, label=None):
        self.label = label or ''

    def __enter__(self):
        self.on = logger.isEnabledFor(logging.INFO)
        if self.on:
            _sync()
        self.start = time.perf_counter()
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        if self.on:
            _sync()
            logger.info(f'{self.label: <40}{(time.perf_counter() - self.start) * 1000:9.3f}ms')
