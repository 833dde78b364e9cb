// vhci_cli.cpp -- the `vhci` executable: the reference's command line
// (vhci -n Ncpus -i input -o freqs [-c checkpoint]) on the device engine.
#include "../../include/vhci.h"

int main(int argc, char** argv) { return vhci_main(argc, argv); }
