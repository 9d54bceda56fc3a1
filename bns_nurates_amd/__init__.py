"""bns_nurates_amd -- MI355X-native gray M1 neutrino opacities."""
