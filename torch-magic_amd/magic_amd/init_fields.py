"""``initialize_fields`` with the reference's name (``init_fields.py:470-532``)."""

from __future__ import annotations

from . import config


def initialize_fields():
    config.default_run().initialize_fields()
