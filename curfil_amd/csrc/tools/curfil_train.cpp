// curfil_train: the command-line tool; its body is cf_tool_train in libcurfil_hip.so.
#include "curfil_hip.h"

int main(int argc, char** argv) { return cf_tool_train(argc, argv); }
