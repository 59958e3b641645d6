"""MI355X-native drop-in for the 2-D grand-potential finite-difference solver loop of MicroSim."""
