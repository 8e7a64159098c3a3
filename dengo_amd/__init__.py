"""MI355X backend for Dengo's generated backward-Euler chemistry/cooling solver."""
from . import chemical_network  # noqa: F401  (registers the ge/de species)
