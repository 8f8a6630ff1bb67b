// cli_exe.cpp -- the roift_oiftrelax executable: roift_cli_main of libroift.so.
#include "../../include/roift.h"

int main(int argc, char **argv) { return roift_cli_main(argc, argv); }
