from .monitors import BunchMonitor
