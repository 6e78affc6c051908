"""MI355X-native residual and implicit-update path of OpenCFD-EC2D (host side)."""
