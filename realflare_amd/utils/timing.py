"""Stage timing in the reference's log format (``realflare/utils/timing.py:8-30``):
``f'{qualname + ":": <40}{ms:9.3f}ms'`` at INFO."""
import logging
import time
from functools import wraps

logger = logging.getLogger(__name__)


def timer(func):
    @wraps(func)
    def wrapper(*args, **kwargs):
        start = time.perf_counter()
        result = func(*args, **kwargs)
        label = f'{func.__qualname__}:'
        logger.info(f'{label: <40}{(time.perf_counter() - start) * 1000:9.3f}ms')
        return result

    return wrapper


class Timer:
    def __init__(self, label=None):
        self.label = label or ''

    def __enter__(self):
        self.start = time.perf_counter()
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        logger.info(f'{self.label: <40}{(time.perf_counter() - self.start) * 1000:9.3f}ms')
