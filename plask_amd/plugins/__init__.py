"""Put this directory on sys.path to import the solvers under the reference's module
names: ``from thermal.static import Static2D``, ``from electrical.shockley import Shockley2D``
(the names PythonManager::loadSolver imports, python/plask/manager.cpp:302-310)."""
import os

PATH = os.path.dirname(os.path.abspath(__file__))
