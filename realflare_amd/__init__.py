"""MI355X-native flare render path with realflare's engine/task interface."""
__version__ = '0.1.0'
