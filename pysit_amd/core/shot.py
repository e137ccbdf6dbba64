"""Shots: a source (or source set) with its receivers (pysit/core/shot.py:12-147)."""

__all__ = ["Shot"]


class Shot(object):
    def __init__(self, sources, receivers):
        self.receivers = receivers
        receivers.set_shot(self)
        self.sources = sources
        sources.set_shot(self)

    def initialize(self, data_length):
        self.receivers.clear_data(data_length)

    def reset_time_series(self, ts):
        self.sources.reset_time_series(ts)
        self.receivers.reset_time_series(ts)

    def gather(self, as_array=False, offset=None):
        if offset is not None:
            raise NotImplementedError("Gather by offset not yet implemented.")
        return self.receivers.data if as_array else self.receivers
