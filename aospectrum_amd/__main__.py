"""``python -m aospectrum_amd band band.toml`` -- the ``aospectrum`` command."""

from .cli import main

raise SystemExit(main())
