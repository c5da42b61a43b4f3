// The device helpers are part of the public plug-in SDK.
#pragma once
#include "../../include/gpuboscia/device/gb_device.h"
